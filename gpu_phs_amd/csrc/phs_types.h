// Data layout shared by the host work-list compiler and the HIP kernels.
//
// A channel (one decay tree) is compiled into a short program of NodeCmd records, one per
// internal branch, in the order the kernel visits them (root first, then a depth-first walk
// that descends into the smaller subtree and defers the other one).  Every record holds all
// constants its branch needs, precomputed on the host, so a wavefront that runs 64 events of
// one channel reads its program with scalar loads and branches only on wave-uniform values.
//
// The random-number index of each branch is fixed by the reference's consumption order, not
// by the visit order: msq randoms in post-order of the internal non-root branches
// (phs_cpu.cpp:180-209), then (phi, cos theta) pairs in pre-order of all internal branches
// (phs_cpu.cpp:301-311).
#pragma once
#include <stdint.h>

namespace gphs {

// Map ids of the reference (phs.h:8-17).
enum MapId : int32_t {
  MAP_NO = 0, MAP_SCHANNEL = 1, MAP_TCHANNEL = 2, MAP_UCHANNEL = 3, MAP_RADIATION = 4,
  MAP_COLLINEAR = 5, MAP_INFRARED = 6, MAP_STEP_E = 11, MAP_STEP_H = 12, MAP_ONSHELL = 99
};

// msq map families as the kernel dispatches them (phs_cpu.cpp:12-61).
enum MsqKind : int32_t { MK_LINEAR = 0, MK_BREIT_WIGNER = 1, MK_EXP = 2, MK_TU = 3, MK_STEP_E = 4, MK_STEP_H = 5 };
// cos-theta map families (phs_cpu.cpp:112-146).
enum CtKind : int32_t { CK_FLAT = 0, CK_PEAKED = 1 };

// One daughter of an internal branch.
struct ChildCmd {
  int32_t leaf;      // 1: final-state particle, 0: internal branch
  int32_t kind;      // MsqKind (internal)
  int32_t xi;        // msq random index (internal)
  int32_t out;       // final-state slot j = log2(bincode) (leaf)
  double msq_leaf;   // m_j^2 (leaf)
  double m_leaf;     // sqrt(m_j^2), rounded as the device would (sqrt is correctly rounded)
  double a0, a1, a2; // msq map constants a[3] (internal), phs_cpu.cpp:452-511
  double m, w;       // map mass and width (internal)
  double msq_min;    // m_min^2, m_min = sum of leaf masses below the branch
  double msq_max;    // m_max^2, m_max = sqrt(s) - m_tot + m_min
  double f_lin;      // (msq_max - msq_min) / s
  double m_min;      // for the per-event sqrt(s) path
  double da_dm[3];   // d a[k] / d M of this daughter's map (adjoint)
  double da_dw[3];   // d a[k] / d Gamma
  int32_t node;      // visit index of this daughter's NodeCmd (internal), -1 for a leaf
  uint32_t leaves;   // bit j set: final-state particle j lies below this daughter (inverse mapping)
  // The daughter's four-momentum as the reference forms it, each branch the sum of its two
  // daughters (phs.cu:1262-1275), as a program of sum_len 4-bit ops (first op in the lowest
  // bits) over two accumulators: j < kSumPush adds final-state momentum j to the current
  // accumulator, kSumPush opens the second one, kSumMerge adds it to the first.  sum_len = 0:
  // the subtree needs more than two accumulators or sixteen ops, and the momenta of `leaves`
  // are added in index order instead.
  int32_t sum_len;
  int32_t pad1;
  uint64_t sum_code;
};

constexpr int kSumPush = 14;
constexpr int kSumMerge = 15;
constexpr int kSumMaxOps = 16;

struct NodeCmd {
  int32_t ct_kind;   // CtKind of this branch's cos-theta map
  int32_t x_phi;     // phi random index; cos theta uses x_phi + 1
  int32_t desc;      // daughter (0/1) visited next, -1: next branch comes from the save bank
  int32_t push;      // daughter (0/1) deferred to the save bank, -1: none
  int32_t map_id;    // reference map id of this branch (for the adjoint / per-event path)
  int32_t parent;    // visit index of the parent branch, -1 for the root
  int32_t slot;      // which daughter (0/1) of the parent this branch is
  int32_t tape_slot; // adjoint tape slot of the boost after this branch, -1: no internal daughter
  int32_t friend_sign; // polar axis: 0 the branch's z axis, +1 / -1 the incoming beam moving along +z / -z
  int32_t pad0;
  double m_max;      // kinematic bound of this branch's mass (root: sqrt(s))
  double b0, b1;     // cos-theta constants b[0], b[1], phs_cpu.cpp:514-527 (friend: at s = 4 E1 E2)
  double m_min;      // sum of leaf masses below the branch
  double m;          // map mass of this branch (cos-theta constants per event)
  double db_dm[2];   // d b[k] / d M (adjoint)
  ChildCmd k[2];
};

struct ChannelDesc {
  int32_t node0;     // first NodeCmd
  int32_t n_nodes;   // n_out - 1
  int32_t n_out;
  int32_t n_x;
  int32_t param0;    // first (M, Gamma) accumulator slot of this channel (adjoint)
  int32_t n_banks;   // save-bank depth the program needs
  int32_t n_tape;    // adjoint tape slots the program needs
  uint32_t div_nx;   // floor(2^32 / n_x) + 1: i / n_x = umulhi(i, div_nx) for i < 2^32 / n_x
  uint32_t div_p;    // floor(2^32 / (4 n_out)) + 1, likewise for the momenta of an event
  int32_t shape;     // entry of the compile-time shape table (phs_shapes.h), -1: none
  double sqrts, s;
  double volume;     // phase-space volume, evaluated in the reference's order
  double bg_lab;     // lab-frame boost beta*gamma = (E1 - E2) / (2 sqrt(E1 E2)), phs.cu:1289
  double g_lab;      // sqrt(1 + bg_lab^2)
  double m_tot;      // sum of all final-state masses
};

// A contiguous run of events of one channel with its device buffers (event-major layouts of
// the reference's interface: x [n][n_x], p [n][n_out][4]).
struct SegDesc {
  int32_t channel;   // index into the plan's channel table
  int32_t align16;   // bit 0: x is 16-byte aligned, bit 1: p is 16-byte aligned
  int64_t n_events;
  const double* x;
  double* p;
  double* factor;
  double* volume;
  uint8_t* ok;
  // adjoint
  const double* g_p;     // [n][n_out][4] cotangents of the momenta, or null
  const double* g_f;     // [n] cotangents of the factor, or null
  double* g_x;           // [n][n_x] gradient w.r.t. the randoms, or null
  double* loss;          // [n] per-event loss (fused loss mode), or null
  // per-event beam energies [n] (phs.cu:1313-1317), or null: the channel's fixed beams
  const double* e1;
  const double* e2;
  // msq of every branch [n][2^n_out - 1] at bincode - 1 (leaves m^2, the root s), as the
  // reference keeps them (phs.cu:570-624), or null; slots of other bincodes are not written
  double* msq;
};

struct TileDesc {
  int32_t seg;
  int32_t count;     // events in this tile
  int64_t e0;        // first event (segment-relative)
};

constexpr int kMaxOut = 12;

}  // namespace gphs
