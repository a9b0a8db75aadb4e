// Forward kernel: the generic variants (programs read from the channel tables) and the
// launch functions.  The device code is in phs_forward.h; the shape-specialised variants are
// instantiated in phs_forward_shapes_*.hip.
#include "phs_forward.h"

namespace gphs {

static_assert(kNumShapes == 64, "the ranges of phs_forward_shapes_*.hip cover the shape table");

namespace {
using namespace fwd;

// Where a plan's randoms live under `lc`: kXLds, or the register row size.
int x_mode(const DevicePlan& dp, const LaunchConfig& lc) {
  if ((lc.lds & 2) || dp.nx_max > 16 || dp.per_event_beams) return kXLds;
  return dp.nx_max <= 8 ? 8 : 16;
}

template <int BANK, int WPE, bool FR>
FwdKernel pick_xn(int xm) {
  if (xm == 8) return forward_kernel<BANK, WPE, 8, FR, false, -1, 0>;
  if (xm == 16) return forward_kernel<BANK, WPE, 16, FR, false, -1, 0>;
  return forward_kernel<BANK, WPE, kXLds, FR, false, -1, 0>;
}

// Plans with friend branches run the variant that has the friend-axis composition, plans with
// per-event beams the one that recomputes the beam-dependent constants (both at the
// compiler-chosen occupancy, the latter with the randoms in LDS); the others carry neither.
template <int BANK>
FwdKernel pick_wpe(const DevicePlan& dp, const LaunchConfig& lc) {
  const int xm = x_mode(dp, lc);
  if (dp.per_event_beams)
    return dp.has_friends ? forward_kernel<BANK, 0, kXLds, true, true, -1, 0>
                           : forward_kernel<BANK, 0, kXLds, false, true, -1, 0>;
  if (dp.has_friends) return pick_xn<BANK, 0, true>(xm);
  if (lc.waves_per_eu == 3) return pick_xn<BANK, 3, false>(xm);
  if (lc.waves_per_eu == 4) return pick_xn<BANK, 4, false>(xm);
  return pick_xn<BANK, 0, false>(xm);
}

// The bank kind a plan runs with under `lc`.
int bank_kind(const DevicePlan& dp, const LaunchConfig& lc) {
  if (dp.n_banks == 0) return BK_NONE;
  if ((lc.lds & 1) || dp.n_banks > 1) return BK_LDS;
  return dp.root_banks_only ? BK_ROOT : BK_VGPR;
}

FwdKernel pick(const DevicePlan& dp, const LaunchConfig& lc) {
  switch (bank_kind(dp, lc)) {
    case BK_NONE: return pick_wpe<BK_NONE>(dp, lc);
    case BK_ROOT: return pick_wpe<BK_ROOT>(dp, lc);
    case BK_VGPR: return pick_wpe<BK_VGPR>(dp, lc);
    default: return pick_wpe<BK_LDS>(dp, lc);
  }
}
}  // namespace

size_t forward_lds_bytes(const DevicePlan& dp, const LaunchConfig& lc) {
  const size_t rows = 4 * (size_t)dp.nout_max + (x_mode(dp, lc) == kXLds ? (size_t)dp.nx_max : 0);
  size_t bytes = rows * (size_t)(lc.block + 1) * sizeof(double);
  if (bank_kind(dp, lc) == BK_LDS) bytes += (size_t)kBankWords * sizeof(double) * lc.block * dp.n_banks;
  return bytes;
}

size_t forward_lds_limit() {
  static size_t limit[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
  if (!limit[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v < 65536) v = 65536;
    limit[dev] = (size_t)std::min(v, 160 * 1024);
  }
  return limit[dev];
}

hipError_t launch_forward_shape(const DevicePlan& dp, int sid, int t_begin, int t_end, const LaunchConfig& lc,
                                hipStream_t stream) {
  if (t_end <= t_begin) return hipSuccess;
  if (sid < 0 || sid >= kNumShapes || lc.block < 64 || lc.block > 256 || lc.block % 64) return hipErrorInvalidValue;
  const int n_out = kShapes[sid].n_out;
  const size_t shmem = 4 * (size_t)n_out * (size_t)(lc.block + 1) * sizeof(double);
  if (shmem > forward_lds_limit()) return hipErrorInvalidValue;
  const int n = t_end - t_begin;
  const int grid = lc.grid > 0 ? (lc.grid < n ? lc.grid : n) : n;
  FwdKernel k = nullptr;
  // one workgroup per tile (grid 0) and tiles of one sub-tile (ept 1): the one-pass kernels
  const bool one = lc.grid == 0 && lc.ept == 1;
  if (dp.per_event_beams) {
    if (sid < kPevShapes) k = pick_shape_pev_0_8(sid, lc.waves_per_eu, lc.block, one);
  } else {
    for (auto pick : {pick_shape_0_8, pick_shape_8_22, pick_shape_22_36, pick_shape_36_50, pick_shape_50_64})
      if (!k) k = pick(sid, lc.waves_per_eu, lc.block, one);
  }
  if (!k) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k, dim3(grid), dim3(lc.block), shmem, stream, dp.tiles, t_begin, t_end, dp.segs, dp.chans,
                     dp.nodes, (int)kShapes[sid].n_x, n_out);
  return hipGetLastError();
}

hipError_t launch_forward(const DevicePlan& dp, int t_begin, int t_end, const LaunchConfig& lc, hipStream_t stream) {
  if (t_end <= t_begin) return hipSuccess;
  if (dp.n_banks < 0 || dp.n_banks > 3 || lc.block < 64 || lc.block > 256 || lc.block % 64) return hipErrorInvalidValue;
  const size_t shmem = forward_lds_bytes(dp, lc);
  if (shmem > forward_lds_limit()) return hipErrorInvalidValue;
  const int n = t_end - t_begin;
  const int grid = lc.grid > 0 ? (lc.grid < n ? lc.grid : n) : n;
  FwdKernel k = pick(dp, lc);
  hipLaunchKernelGGL(k, dim3(grid), dim3(lc.block), shmem, stream, dp.tiles, t_begin, t_end, dp.segs, dp.chans,
                     dp.nodes, dp.nx_max, dp.nout_max);
  return hipGetLastError();
}

}  // namespace gphs
