// Launch interface between the host library and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "phs_types.h"

namespace gphs {

// Device copies of one plan's tables.
struct DevicePlan {
  const TileDesc* tiles = nullptr;
  const SegDesc* segs = nullptr;
  const ChannelDesc* chans = nullptr;
  const NodeCmd* nodes = nullptr;
  int n_tiles = 0;
  int n_banks = 0;       // save-bank depth the deepest program needs (0..3)
  int nx_max = 0;        // largest n_x of the plan's channels
  int nout_max = 0;      // largest n_out of the plan's channels
  int root_banks_only = 0;  // 1: no branch but a root defers a subtree, and the root has no friend
  int has_friends = 0;   // 1: some branch takes its polar axis from an incoming beam
  int per_event_beams = 0;  // 1: some segment carries per-event beam energies
};

struct LaunchConfig {
  int block = 256;       // threads per workgroup (multiple of 64, <= 256)
  int ept = 1;           // events per thread per tile
  int grid = 0;          // workgroups; 0 = one per tile, else a persistent grid that strides
  int waves_per_eu = 0;  // register-allocation occupancy target: 0 (compiler's choice) or 3
  int lds = 0;           // bit 0: save banks of deferred subtrees in LDS (always for 2+ levels),
                         // bit 1: each event's randoms in LDS instead of registers
};

// LDS bytes per workgroup of the forward kernel under `lc`: the staged randoms and momenta of
// `lc.block` events, plus the save banks when they are kept in LDS.
size_t forward_lds_bytes(const DevicePlan& dp, const LaunchConfig& lc);
// LDS a workgroup may allocate on the current device.
size_t forward_lds_limit();

// Forward pass over tiles [t_begin, t_end).
hipError_t launch_forward(const DevicePlan& dp, int t_begin, int t_end, const LaunchConfig& lc, hipStream_t stream);
// The same for tiles whose channels all have tree shape `sid` (phs_shapes.h), with the kernel
// specialised for it (no friends; lc.lds is not used: randoms in registers).  With per-event
// beams (dp.per_event_beams) only shapes [0, kPevShapes) -- the trees of up to four final-state
// particles -- have such kernels.
constexpr int kPevShapes = 8;
// The adjoint's specialised shapes: [0, kAdjShapes), the trees of up to four final-state
// particles (the others run the generic adjoint).
constexpr int kAdjShapes = 8;
hipError_t launch_forward_shape(const DevicePlan& dp, int sid, int t_begin, int t_end, const LaunchConfig& lc,
                                hipStream_t stream);

// Per-lane LDS layout of the adjoint (in doubles, lane-strided): boost parameters of every
// branch [0, off_in), branch inputs [off_in, off_st), stash [off_st, off_acc), (M, Gamma)
// accumulators [off_acc, words).
struct AdjLayout {
  int off_in, off_st, off_acc, words;
  int compact;   // 1: kAdjCompactStashWords per stash level (only the root defers), else kAdjStashWords
};
constexpr int kAdjStashWords = 19, kAdjCompactStashWords = 15;

// Adjoint over all tiles, one wavefront per 64-event sub-tile (n_sub per tile), then the
// per-channel reduction of the parameter partials (through granges, kAdjRanges * part_stride
// doubles per channel) into gparam[2 * (chans[c].param0 + node) + {0: M, 1: Gamma}].  done:
// n_chans counters that are 0 before the launch (and are 0 again after it).
//
// The tiles go in runs of [begin, end) (together every tile once), each run on the kernel
// specialised for tree shape `sid` when there is one (phs_shapes.h), else the generic kernel.
struct TileRun {
  int sid, begin, end;
};
constexpr int kAdjRanges = 16;
hipError_t launch_adjoint(const DevicePlan& dp, const AdjLayout& lay, const TileRun* runs, int n_runs, int loss_mode,
                          double* gpart, int part_stride, int n_sub, double* granges, unsigned int* done,
                          const int* chan_tiles, const int* chan_tile_off, int n_chans, double* gparam,
                          hipStream_t stream);

// Inverse mapping.  Per segment of the plan: the channel that generated its events (its
// factor and randoms are copied) and the output row of its first event.
struct InvSeg {
  int32_t own;
  int32_t pad;
  int64_t row0;
};
// One channel of the inverse's channel order: channels whose roots share their evaluation
// (the same daughters' final-state particles, added in the same association) follow one
// another, and the first of each group has new_group set.
struct InvStep {
  int32_t channel;
  int32_t new_group;
};
// Device copies of one process's complete channel table (every channel is a target).
struct InverseTables {
  const InvSeg* isegs = nullptr;
  const ChannelDesc* chans = nullptr;
  const NodeCmd* nodes = nullptr;
  const InvStep* steps = nullptr;
  int n_ch = 0, n_out = 0, n_x = 0, n_banks = 0, has_friends = 0, per_event_beams = 0;
  int root_bank = 0;   // every deferred subtree is a daughter of the root and no root has a friend
  // shared roots (no friends): the channels in the order of `steps`; save-bank levels of the
  // subtrees deferred below the root
  int shared_roots = 0, rest_banks = 0;
  int generic = 0;   // some channel's program has none of the shapes of up to four final-state particles
};
// The inverse walks specialised per tree shape cover the shapes 0 .. kInvShapes - 1 of the
// compile-time table (the trees of two to four final-state particles).
constexpr int kInvShapes = 8;
size_t inverse_lds_bytes(const InverseTables& it, int block);
// factors [n_rows][n_ch] and x_out [n_ch][n_rows][n_x] for the events of tiles [0, n_tiles).
hipError_t launch_inverse(const InverseTables& it, const TileDesc* tiles, int n_tiles, const SegDesc* segs,
                          int block, int grid, double* factors, double* x_out, int64_t n_rows, hipStream_t stream);

// Adaptive grids of the randoms (rng.cpp:5: 10 equal bins per channel and dimension).
constexpr int kGridBins = 10;
// Randoms of every segment of the plan drawn from the grids: cdf [plan channel][n_dim][11].
hipError_t launch_grid_generate(const DevicePlan& dp, const double* cdf, int n_dim, uint64_t seed, hipStream_t stream);
// counts [plan channel][n_dim][10] histogram of the accepted events' randoms, then the accepted
// count of every plan channel (n_chans more entries); cleared first.
hipError_t launch_grid_histogram(const DevicePlan& dp, int n_chans, int n_dim, unsigned long long* counts,
                                 hipStream_t stream);

}  // namespace gphs
