// Adjoint kernel: per event, the forward pass with a tape, then the reverse sweep
// (phs_adjoint.h).  One workgroup = one wavefront = 64 events of one tile; the tape, the
// stash and the lane's (M, Gamma) accumulators live in LDS, lane-strided so every access is
// conflict-free.  The parameter adjoints of a tile are reduced across the wavefront and
// written as one partial per tile; a second kernel sums the partials of each channel in a
// fixed order, so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <utility>

#include "phs_adjoint.h"
#include "phs_kernels.h"
#include "phs_shapes.h"

namespace gphs {

namespace {

// COMPACT: the stash keeps 15 of its 19 words (only the root defers; phs_adjoint.h).  Word w
// of a lane is at base[w * 64] (one wavefront per workgroup: consecutive lanes, consecutive
// words).
template <bool COMPACT>
struct LdsStore {
  static constexpr int stride = 64;
  double* base;
  int off_in, off_st;
  __device__ double& T(int node, int k) { return base[(node * 6 + k) * stride]; }
  __device__ double& in(int node, int k) { return base[(off_in + node * 2 + k) * stride]; }
  __device__ __forceinline__ void st_put(int level, int k, double v) {
    if constexpr (COMPACT) {
      const int c = compact_stash_slot(k);
      if (c >= 0) base[(off_st + level * kAdjCompactStashWords + c) * stride] = v;
    } else {
      base[(off_st + level * kAdjStashWords + k) * stride] = v;
    }
  }
  __device__ __forceinline__ double st_get(int level, int k) const {
    if constexpr (COMPACT) {
      const int c = compact_stash_slot(k);
      return c >= 0 ? base[(off_st + level * kAdjCompactStashWords + c) * stride] : 0.0;
    } else {
      return base[(off_st + level * kAdjStashWords + k) * stride];
    }
  }
};

struct LdsAcc {
  double* base;
  __device__ double& operator[](int i) { return base[i * 64]; }
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int kAdjXReg = 8;             // n_x <= 8 <=> n_out <= 4
constexpr int kAdjMultiShapes = kAdjShapes;   // shapes 0-7 of the table: the trees of 2-4 final-state particles
constexpr int kSmallParams = 2 * 3;     // (M, Gamma) of the 3 branches of an n_out = 4 tree

// One 64-event unit of the adjoint: sub-tile `sub` of tile t, events [64 sub, 64 sub + 64).
// XN > 0: every channel has at most XN randoms, and each event's row is held in registers.
// SID >= 0: the tile's channel has that compile-time tree shape (XN is then its n_x).
// Partials go to the tile's global unit.
template <int XN, bool COMPACT, int SID, int LM, bool PEV, bool FR>
__device__ __forceinline__ void adjoint_unit(int t, int sub, int n_sub, const TileDesc& td, const SegDesc& sd,
                                             const ChannelDesc& cd, const NodeCmd* __restrict__ prog,
                                             double* __restrict__ gpart, int part_stride, const AdjLayout& lay,
                                             double* lds) {
  const int lane = threadIdx.x;
  LdsStore<COMPACT> S{lds + lane, lay.off_in, lay.off_st};
  // the (M, Gamma) accumulators: in registers for the small programs (XN > 0: n_out <= 4, at
  // most kSmallParams), else in LDS after the tape
  double accv[XN > 0 ? kSmallParams : 1];
  LdsAcc acc_lds{lds + lay.off_acc * 64 + lane};
  const int np = 2 * cd.n_nodes;
  if constexpr (XN > 0) {
#pragma unroll
    for (int v = 0; v < kSmallParams; v++) accv[v] = 0;
  } else {
    for (int v = 0; v < np; v++) acc_lds[v] = 0;
  }
  const int end = min(td.count, 64 * (sub + 1));
  // a unit has at most 64 events: one per lane, no loop (constants are not hoisted out of one)
  if (const int local = 64 * sub + lane; local < end) {
    const int64_t e = td.e0 + local;
    const double* xg = sd.x + e * cd.n_x;
    const double* gp = sd.g_p ? sd.g_p + e * (int64_t)cd.n_out * 4 : nullptr;
    const double gf = sd.g_f ? sd.g_f[e] : 0.0;
    double* p_out = sd.p ? sd.p + e * (int64_t)cd.n_out * 4 : nullptr;
    double F = 0, L = 0;
    bool ok;
    EvBeams eb;
    if constexpr (PEV) eb = event_beams(cd, sd.e1[e], sd.e2[e]);
    else eb = fixed_beams(cd);
    if constexpr (XN > 0) {
      double xv[XN];
#pragma unroll
      for (int k = 0; k < XN; k++) xv[k] = xg[SID >= 0 || k < cd.n_x ? k : cd.n_x - 1];   // entries past n_x are never read
      ok = event_adjoint<SID, LM, PEV, FR>(cd, eb, prog, xv, gp, gf, p_out, &F, &L, sd.g_x + e * cd.n_x, accv, S);
    } else {
      ok = event_adjoint<SID, LM, PEV, FR>(cd, eb, prog, xg, gp, gf, p_out, &F, &L, sd.g_x + e * cd.n_x, acc_lds, S);
    }
    if (sd.factor) sd.factor[e] = F;
    if (sd.volume) sd.volume[e] = eb.volume;
    if (sd.ok) sd.ok[e] = ok ? 1 : 0;
    if (sd.loss) sd.loss[e] = L;
  }
  if constexpr (XN > 0) {
#pragma unroll
    for (int v = 0; v < kSmallParams; v++) {
      if (v < np) {   // uniform
        const double tot = wave_sum(accv[v]);
        if (lane == 0) gpart[((int64_t)t * n_sub + sub) * part_stride + v] = tot;
      }
    }
  } else {
    for (int v = 0; v < np; v++) {
      const double tot = wave_sum(acc_lds[v]);
      if (lane == 0) gpart[((int64_t)t * n_sub + sub) * part_stride + v] = tot;
    }
  }
}

// The launch covers tiles [t_begin, t_begin + n_units / n_sub): unit u is sub-tile (u % n_sub)
// of tile t_begin + u / n_sub.  SID >= 0: every tile's channel has that shape; SID = kAdjMulti:
// every tile's channel has one of the shapes of up to four final-state particles, and each
// workgroup takes the specialised code of its tile's shape (one launch for several shapes:
// a launch lasts at least one wave's lifetime, so a launch per shape would add that much per
// shape).  The register budget keeps two waves per SIMD (at most 256 VGPRs).  PEV: per-event
// beam energies (loss mode 0).  FR: plans with friend branches (the generic walk, randoms read
// from memory, accumulators in LDS).
constexpr int kAdjMulti = 1000;
template <int XN, bool COMPACT, int SID, int LM, bool PEV = false, bool FR = false>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void adjoint_kernel(
    const TileDesc* __restrict__ tiles, int t_begin, int n_units, int n_sub, const SegDesc* __restrict__ segs,
    const ChannelDesc* __restrict__ chans, const NodeCmd* __restrict__ nodes, double* __restrict__ gpart,
    int part_stride, AdjLayout lay) {
  extern __shared__ double lds[];
  // one workgroup per unit (the host launches n_units workgroups): no loop
  if (const int u = (int)blockIdx.x; u < n_units) {
    const int tl = u / n_sub, sub = u - tl * n_sub, t = t_begin + tl;
    const TileDesc td = tiles[t];
    const SegDesc sd = segs[td.seg];
    const ChannelDesc cd = chans[sd.channel];
    const NodeCmd* prog = nodes + cd.node0;
    if constexpr (SID == kAdjMulti) {
      switch (cd.shape) {
        case 0: adjoint_unit<kShapes[0].n_x, COMPACT, 0, LM, PEV, false>(t, sub, n_sub, td, sd, cd, prog, gpart, part_stride, lay, lds); break;
        case 1: adjoint_unit<kShapes[1].n_x, COMPACT, 1, LM, PEV, false>(t, sub, n_sub, td, sd, cd, prog, gpart, part_stride, lay, lds); break;
        case 2: adjoint_unit<kShapes[2].n_x, COMPACT, 2, LM, PEV, false>(t, sub, n_sub, td, sd, cd, prog, gpart, part_stride, lay, lds); break;
        case 3: adjoint_unit<kShapes[3].n_x, COMPACT, 3, LM, PEV, false>(t, sub, n_sub, td, sd, cd, prog, gpart, part_stride, lay, lds); break;
        case 4: adjoint_unit<kShapes[4].n_x, COMPACT, 4, LM, PEV, false>(t, sub, n_sub, td, sd, cd, prog, gpart, part_stride, lay, lds); break;
        case 5: adjoint_unit<kShapes[5].n_x, COMPACT, 5, LM, PEV, false>(t, sub, n_sub, td, sd, cd, prog, gpart, part_stride, lay, lds); break;
        case 6: adjoint_unit<kShapes[6].n_x, COMPACT, 6, LM, PEV, false>(t, sub, n_sub, td, sd, cd, prog, gpart, part_stride, lay, lds); break;
        case 7: adjoint_unit<kShapes[7].n_x, COMPACT, 7, LM, PEV, false>(t, sub, n_sub, td, sd, cd, prog, gpart, part_stride, lay, lds); break;
        default: break;   // not reached: the host launches only tiles of shapes 0-7 here
      }
    } else {
      adjoint_unit<XN, COMPACT, SID, LM, PEV, FR>(t, sub, n_sub, td, sd, cd, prog, gpart, part_stride, lay, lds);
    }
  }
}

// gparam[2 * slot + w] = sum over the sub-tiles of the tiles of the slot's channel (every slot
// is written, channels without events get 0).  kSplit workgroups per channel each sum a
// contiguous range of the channel's units (thread t adds units t, t + 256, ..., then a fixed
// tree: a butterfly within each wave, the four waves in order), and the workgroup that finishes
// its channel's last range adds the kSplit range sums in order.  Every sum has a fixed order,
// so results do not depend on scheduling.  `done` counts the finished ranges per channel; the
// last workgroup resets it to 0, so it is 0 again when the launch ends.
constexpr int kReduceThreads = 256;
constexpr int kSplit = kAdjRanges;
constexpr int kMaxParams = 2 * (kMaxOut - 1);
__global__ __launch_bounds__(kReduceThreads) void param_reduce(
    const double* __restrict__ gpart, int part_stride, int n_sub, const int* __restrict__ chan_tiles,
    const int* __restrict__ chan_tile_off, const ChannelDesc* __restrict__ chans, int n_chans,
    double* __restrict__ granges, unsigned int* __restrict__ done, double* __restrict__ gparam) {
  __shared__ double red[kReduceThreads / 64][kMaxParams];
  __shared__ int last;
  const int c = blockIdx.x / kSplit, r = blockIdx.x % kSplit;
  if (c >= n_chans) return;
  const int np = 2 * chans[c].n_nodes;
  const int t0 = chan_tile_off[c];
  const int n_units = (chan_tile_off[c + 1] - t0) * n_sub;
  const int chunk = (n_units + kSplit - 1) / kSplit;
  const int u0 = r * chunk, u1 = min(n_units, u0 + chunk);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  double sum[kMaxParams];
#pragma unroll
  for (int v = 0; v < kMaxParams; v++) sum[v] = 0;
  for (int k = u0 + tid; k < u1; k += kReduceThreads) {
    const int tile = chan_tiles[t0 + k / n_sub];
    const double* src = gpart + ((int64_t)tile * n_sub + k % n_sub) * part_stride;
#pragma unroll
    for (int v = 0; v < kMaxParams; v++)
      if (v < np) sum[v] += src[v];
  }
#pragma unroll
  for (int v = 0; v < kMaxParams; v++) {
    if (v < np) {   // uniform
      const double t = wave_sum(sum[v]);
      if (lane == 0) red[wv][v] = t;
    }
  }
  __syncthreads();
  if (tid < np) granges[(int64_t)blockIdx.x * part_stride + tid] = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
  // publish this range, then count it; the workgroup that counts the channel's last range
  // sees every range (device-scope release by the writers, acquire by the last one)
  __threadfence();
  __syncthreads();
  if (tid == 0) last = atomicAdd(&done[c], 1u) == (unsigned int)(kSplit - 1);
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (tid < np) {
    double s = 0;
    for (int q = 0; q < kSplit; q++) s += granges[((int64_t)c * kSplit + q) * part_stride + tid];
    gparam[2 * chans[c].param0 + tid] = s;
  }
  if (tid == 0) done[c] = 0;
}

using AdjKernel = void (*)(const TileDesc*, int, int, int, const SegDesc*, const ChannelDesc*, const NodeCmd*,
                           double*, int, AdjLayout);

// The adjoint specialised for a tree shape of up to four final-state particles (n_x <= 8): the
// randoms in the shape's n_x registers, the accumulators in registers; larger shapes run the
// generic kernel.
template <int SID, int LM>
AdjKernel adj_shape(bool compact) {
  if constexpr (kShapes[SID].n_out <= 4) {
    constexpr int xn = kShapes[SID].n_x;
    return compact ? adjoint_kernel<xn, true, SID, LM> : adjoint_kernel<xn, false, SID, LM>;
  } else {
    return nullptr;
  }
}
template <int... S>
AdjKernel adj_pick_seq(int sid, bool compact, int lm, std::integer_sequence<int, S...>) {
  AdjKernel k = nullptr;
  (void)((sid == S ? (k = lm ? adj_shape<S, 1>(compact) : adj_shape<S, 0>(compact), true) : false) || ...);
  return k;
}
template <int LM>
AdjKernel adj_generic(const DevicePlan& dp, const AdjLayout& lay) {
  if (dp.has_friends) {   // the host sets no compact stash and keeps the accumulators in LDS
    if constexpr (LM == 0) {
      if (dp.per_event_beams) return adjoint_kernel<0, false, -1, 0, true, true>;
    }
    return adjoint_kernel<0, false, -1, LM, false, true>;
  }
  if constexpr (LM == 0) {   // per-event beams: the external cotangents only (the built-in loss assumes equal beams)
    if (dp.per_event_beams)
      return dp.nx_max > kAdjXReg ? adjoint_kernel<0, false, -1, 0, true>
             : lay.compact ? adjoint_kernel<kAdjXReg, true, -1, 0, true> : adjoint_kernel<kAdjXReg, false, -1, 0, true>;
  }
  return dp.nx_max > kAdjXReg ? adjoint_kernel<0, false, -1, LM>
         : lay.compact ? adjoint_kernel<kAdjXReg, true, -1, LM> : adjoint_kernel<kAdjXReg, false, -1, LM>;
}

}  // namespace

hipError_t launch_adjoint(const DevicePlan& dp, const AdjLayout& lay, const TileRun* runs, int n_runs, int loss_mode,
                          double* gpart, int part_stride, int n_sub, double* granges, unsigned int* done,
                          const int* chan_tiles, const int* chan_tile_off, int n_chans, double* gparam,
                          hipStream_t stream) {
  const size_t shmem = (size_t)lay.words * 64 * sizeof(double);
  if (shmem > forward_lds_limit() || n_sub < 1) return hipErrorInvalidValue;
  if ((int64_t)dp.n_tiles * n_sub > 0x7fffffff) return hipErrorInvalidValue;
  if (dp.per_event_beams && loss_mode != 0) return hipErrorInvalidValue;
  AdjKernel generic = loss_mode ? adj_generic<1>(dp, lay) : adj_generic<0>(dp, lay);
  AdjKernel multi = loss_mode ? (lay.compact ? adjoint_kernel<kAdjXReg, true, kAdjMulti, 1>
                                             : adjoint_kernel<kAdjXReg, false, kAdjMulti, 1>)
                              : (lay.compact ? adjoint_kernel<kAdjXReg, true, kAdjMulti, 0>
                                             : adjoint_kernel<kAdjXReg, false, kAdjMulti, 0>);
  // per-event beams: the shapes of up to four final-state particles take the multi-shape
  // kernel also for a run of one shape (one kernel, not one per shape)
  AdjKernel multi_pev = lay.compact ? adjoint_kernel<kAdjXReg, true, kAdjMulti, 0, true>
                                    : adjoint_kernel<kAdjXReg, false, kAdjMulti, 0, true>;
  auto small = [](int sid) { return sid >= 0 && sid < kAdjMultiShapes; };
  for (int r = 0; r < n_runs;) {
    // consecutive runs of the shapes of up to four final-state particles: one launch
    int r1 = r + 1, end = runs[r].end;
    if (small(runs[r].sid))
      while (r1 < n_runs && small(runs[r1].sid) && runs[r1].begin == end) end = runs[r1++].end;
    const int units = (end - runs[r].begin) * n_sub;
    AdjKernel k = nullptr;
    if (dp.has_friends) k = generic;
    else if (dp.per_event_beams) k = small(runs[r].sid) ? multi_pev : generic;
    else if (r1 - r > 1) k = multi;
    else if (runs[r].sid >= 0)
      k = adj_pick_seq(runs[r].sid, lay.compact != 0, loss_mode, std::make_integer_sequence<int, kNumShapes>{});
    if (!k) k = generic;
    if (units > 0) {
      hipLaunchKernelGGL(k, dim3(units), dim3(64), shmem, stream, dp.tiles, runs[r].begin, units, n_sub, dp.segs,
                         dp.chans, dp.nodes, gpart, part_stride, lay);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
    r = r1;
  }
  if (n_chans > 0) {
    hipLaunchKernelGGL(param_reduce, dim3(n_chans * kSplit), dim3(kReduceThreads), 0, stream, gpart, part_stride,
                       n_sub, chan_tiles, chan_tile_off, dp.chans, n_chans, granges, done, gparam);
  }
  return hipGetLastError();
}

}  // namespace gphs
