// Host work-list compiler: a process in the reference's tree format -> per-channel kernel
// programs (phs_types.h).  Replaces the reference's global tables and its command-list
// extraction (phs.cu:248-314, 327-509; global_phs.cpp:6-27) with one self-contained,
// per-process object.
#pragma once
#include <string>
#include <vector>

#include "phs_types.h"

namespace gphs {

struct ProcessSpec {
  int n_in = 2, n_out = 0, n_channels = 0;
  // Reference format (whizard_gpu_phs.cpp:57-87): daughters as bincodes (0 = none),
  // [n_channels][n_prt]; maps [n_channels][n_prt_out], index = bincode - 1.
  std::vector<int> d1, d2, has_children, friends;
  std::vector<int> map_ids;
  std::vector<double> map_masses, map_widths;
  std::vector<double> flv_masses, flv_widths;   // [n_in + n_out], incoming first
  double E1 = 0, E2 = 0;                          // beam energies

  int n_prt() const { return (1 << (n_in + n_out)) - 1; }
  int n_prt_out() const { return (1 << n_out) - 1; }
  int n_x() const { return 3 * n_out - 4; }
  int root() const { return n_prt_out() - 1; }
};

struct CompiledProcess {
  ProcessSpec spec;
  double sqrts = 0;
  std::vector<ChannelDesc> channels;   // node0 relative to `nodes`
  std::vector<NodeCmd> nodes;
  std::vector<int> node_branch;        // branch index (bincode - 1) of every NodeCmd
  std::vector<double> mass_sum;        // [n_channels][n_prt_out]
  int max_banks = 0;
  bool has_friends = false;            // some branch takes its polar axis from an incoming beam
};

// Validate and compile.  Returns an empty string on success, else the reason.
std::string compile_process(const ProcessSpec& spec, CompiledProcess& out);

// Volume of one channel at sqrt(s), in the reference's evaluation order
// (phs_cpu.cpp:195-212 recursion).
double channel_volume(const ProcessSpec& spec, int c, double sqrts);

}  // namespace gphs
