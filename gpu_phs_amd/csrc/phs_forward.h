// Forward phase-space kernel (device code shared by the translation units that instantiate
// it: phs_forward.hip for the generic variants, phs_forward_shapes_*.hip for the
// shape-specialised ones): msq mapping -> angles -> boost chain, fused, one lane per
// event, all per-event state in VGPRs.
//
// Replaces the reference's pipeline of 3 + 2(n-1) + 2 launches with HBM-resident scratch
// (_init_*, _apply_msq, _create_boosts_step_*, _apply_boost_targets, _boost_to_lab_frame;
// phs.cu:1430-1560).  A workgroup runs one tile (a contiguous run of events of one channel);
// the channel's program (NodeCmd records, phs_types.h) is read with scalar loads, so map
// selection is a wave-uniform branch instead of the reference's per-lane function pointers
// (mappings_gpu.hpp:312-363).  The boost matrix of the branch being visited stays in
// registers; the matrix of a deferred subtree waits in a save bank (registers or LDS).
//
// Each lane loads its event's randoms (one row of [n][n_x]) when the event starts, so the
// branch loop never waits on memory: into 8 or 16 registers, indexed with the program's
// wave-uniform random indices (register indexing by a scalar), or into the lane's column of an
// LDS table (always for more than 16 randoms, n_out > 6).  The momenta, the bulk of the
// traffic, go through LDS: each lane writes its event's four-vectors into a table
// p[4 j + mu][event] (consecutive lanes, consecutive words: conflict-free), and the workgroup
// then writes the sub-tile's momenta ([n][n_out][4]) as one contiguous run of 16-byte words,
// so the stores leave the CU as whole cache lines instead of 64 scattered 16-byte pieces.
// Rows are padded to block + 1 doubles so that the gather of the store phase spreads over the
// LDS banks.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "phs_device.h"
#include "phs_kernels.h"
#include "phs_shapes.h"
#include "phs_walk.h"

namespace gphs {

namespace fwd {


// Save banks: where the walk keeps what it needs to resume a deferred subtree -- the boost M
// after the branch that deferred it and the subtree's incoming msq, m and p_decay.
//   BK_NONE  no branch defers a subtree (every branch has at most one internal daughter);
//   BK_ROOT  only the root defers: its boost is L1 itself, kept as its six parameters and
//            re-formed at the pop (9 doubles of registers instead of 19);
//   BK_VGPR  one level anywhere in the tree, in registers;
//   BK_LDS   any number of levels in LDS.
enum BankKind { BK_NONE = 0, BK_ROOT = 1, BK_VGPR = 2, BK_LDS = 3 };

struct L1Par { double g, bg, ct, st, cp, sp; };

template <int KIND>
struct Bank {   // BK_NONE
  __device__ __forceinline__ void push(int, const double (&)[16], const L1Par&, double, double, double) {}
  __device__ __forceinline__ void pop(int, double (&)[16], double&, double&, double&) const {}
};

template <>
struct Bank<BK_ROOT> {
  L1Par l;
  double msq, m, p;
  __device__ __forceinline__ void push(int, const double (&)[16], const L1Par& par, double m2, double mm, double pp) {
    l = par; msq = m2; m = mm; p = pp;
  }
  __device__ __forceinline__ void pop(int, double (&M)[16], double& m2, double& mm, double& pp) const {
    set_l1(M, l.g, l.bg, l.ct, l.st, l.cp, l.sp);
    m2 = msq; mm = m; pp = p;
  }
};

template <>
struct Bank<BK_VGPR> {
  double L[16];
  double msq, m, p;
  __device__ __forceinline__ void push(int, const double (&M)[16], const L1Par&, double m2, double mm, double pp) {
#pragma unroll
    for (int i = 0; i < 16; i++) L[i] = M[i];
    msq = m2; m = mm; p = pp;
  }
  __device__ __forceinline__ void pop(int, double (&M)[16], double& m2, double& mm, double& pp) const {
#pragma unroll
    for (int i = 0; i < 16; i++) M[i] = L[i];
    m2 = msq; mm = m; pp = p;
  }
};

// In LDS: level d of lane t holds kBankWords doubles at base[(kBankWords d + i) * block + t]
// (consecutive lanes hit consecutive 8-byte words, so every access is conflict-free).
constexpr int kBankWords = 19;
template <>
struct Bank<BK_LDS> {
  double* base;
  int stride;
  __device__ __forceinline__ void push(int d, const double (&M)[16], const L1Par&, double m2, double mm, double pp) {
    double* b = base + kBankWords * d * stride;
#pragma unroll
    for (int i = 0; i < 16; i++) b[i * stride] = M[i];
    b[16 * stride] = m2;
    b[17 * stride] = mm;
    b[18 * stride] = pp;
  }
  __device__ __forceinline__ void pop(int d, double (&M)[16], double& m2, double& mm, double& pp) const {
    const double* b = base + kBankWords * d * stride;
#pragma unroll
    for (int i = 0; i < 16; i++) M[i] = b[i * stride];
    m2 = b[16 * stride];
    mm = b[17 * stride];
    pp = b[18 * stride];
  }
};

template <int KIND>
__device__ __forceinline__ Bank<KIND> make_bank(double* bank_base) {
  if constexpr (KIND == BK_LDS) return Bank<BK_LDS>{bank_base + threadIdx.x, (int)blockDim.x};
  else return Bank<KIND>{};
}

extern __shared__ double lds[];


// Randoms of one event: with XN > 0 the row is held in XN registers (n_x <= XN) and indexed
// by the program's wave-uniform indices; with XN = kXLds it is copied into the lane's column
// of an LDS table x[k][event] when the event starts.
constexpr int kXLds = -1;

// One event of channel `cd` (program `prog`).  xv / xl: the event's randoms (registers, or the
// lane's LDS column); pl: this lane's column of the momentum table (word 4 j + mu at
// pl[(4 j + mu) * S]).
// eb: the event's beams (the channel's fixed ones unless PEV: then the map constants that
// depend on sqrt(s) are recomputed per event as well).  msq_row: this event's row of the
// per-branch msq output when want_msq (a wave-uniform flag), else unused.
// The randoms of event row `xg` into registers (XN > 0: entries past n_x repeat the last one and
// are never read; a shape kernel's XN is its n_x) or into the lane's LDS column xl (XN = kXLds).
template <int XN, int SID>
__device__ __forceinline__ void load_randoms(const double* __restrict__ xg, int n_x, double (&xv)[XN > 0 ? XN : 1],
                                             double* __restrict__ xl, int S) {
  const gdouble* __restrict__ xrow = as_global(xg);
  if constexpr (SID >= 0) {
#pragma unroll
    for (int k = 0; k < XN; k++) xv[k] = xrow[k];
  } else if constexpr (XN > 0) {
#pragma unroll
    for (int k = 0; k < XN; k++) xv[k] = xrow[k < n_x ? k : n_x - 1];
  } else {
    for (int k = 0; k < n_x; k++) xl[k * S] = xrow[k];
  }
}

template <int BANK, int XN, bool FR, bool PEV, int SID>
__device__ __forceinline__ void forward_event(const ChannelDesc& cd, const EvBeams& eb, const NodeCmd* __restrict__ prog,
                                              const double (&xv)[XN > 0 ? XN : 1], double* __restrict__ pl,
                                              const double* __restrict__ xl, int S, double* bank_base, bool want_msq,
                                              gdouble* __restrict__ msq_row, double& factor_out, bool& ok_out) {
  using W = Walk<SID>;
  const double s = eb.s;
#define GPHS_X(k) (XN > 0 ? xv[XN > 0 ? (k) : 0] : xl[(k) * S])

  // boost of the parent branch; the root's parent is the identity (phs_cpu.cpp:568-573)
  double M[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  auto bank = make_bank<BANK>(bank_base);
  double cur_msq = s;    // root: msq = s (phs_cpu.cpp:195)
  double cur_p = 0.0;    // root: p_decay = 0 (never set by the reference)
  double factor = 1.0;
  bool ok = true;
  int depth = 0;
  double cur_m = eb.sqrts;  // sqrt(msq) of the branch being visited, carried from its parent
  if (want_msq) msq_row[(1 << cd.n_out) - 2] = s;

  auto branch = [&](const int i) {
    const NodeCmd& nc = prog[i];
    const bool leaf0 = W::leaf(nc, i, 0), leaf1 = W::leaf(nc, i, 1);
    double m_max = nc.m_max, cb0 = nc.b0, cb1 = nc.b1;
    if constexpr (PEV) node_at(nc, i == 0, eb, cd.m_tot, m_max, cb0, cb1);
    // daughters' msq (phs_cpu.cpp:180-191, 207-208; leaves phs_cpu.cpp:541-547) and masses
    double msq0, f0, m1, msq1, f1, m2;
    if (leaf0) { msq0 = nc.k[0].msq_leaf; f0 = 1.0; m1 = nc.k[0].m_leaf; }
    else if constexpr (PEV) {
      msq_map(child_at(nc.k[0], eb.sqrts, s, cd.m_tot), GPHS_X(W::xi(nc, i, 0)), s, msq0, f0); m1 = math::sqrt_phys(msq0);
    } else { msq_map(nc.k[0], GPHS_X(W::xi(nc, i, 0)), s, msq0, f0); m1 = math::sqrt_phys(msq0); }
    if (leaf1) { msq1 = nc.k[1].msq_leaf; f1 = 1.0; m2 = nc.k[1].m_leaf; }
    else if constexpr (PEV) {
      msq_map(child_at(nc.k[1], eb.sqrts, s, cd.m_tot), GPHS_X(W::xi(nc, i, 1)), s, msq1, f1); m2 = math::sqrt_phys(msq1);
    } else { msq_map(nc.k[1], GPHS_X(W::xi(nc, i, 1)), s, msq1, f1); m2 = math::sqrt_phys(msq1); }
    if (want_msq) {
      msq_row[nc.k[0].leaves - 1] = msq0;
      msq_row[nc.k[1].leaves - 1] = msq1;
    }

    // 1 -> 2 decay of this branch (phs_cpu.cpp:218-233)
    const double m = cur_m;
    const double lda = (cur_msq - msq0 - msq1) * (cur_msq - msq0 - msq1) - 4 * msq0 * msq1;
    ok = ok && (lda > 0 && m > m1 + m2 && m <= m_max);
    // one reciprocal of m (estimate + two Newton steps) serves p = sqrt(lambda) / (2 m) and
    // bg = p_parent / m; the results differ from the quotients by an ulp or so
    const double sl = math::sqrt_phys(lda);
    const double rm = math::recip(m);
    const double pk = 0.5 * (sl * rm);
    factor *= f0 * f1 * math::quot(sl, cur_msq);

    // angles and the boost of this branch (phs_cpu.cpp:299-340)
    const double bg = m > 0 ? cur_p * rm : 0.0;
    const double g = math::sqrt_phys(1 + bg * bg);
    double sp, cp;
    math::sincos_turns(GPHS_X(W::xphi(nc, i)), sp, cp);  // phi = 2 pi x
    double ct, st, fct;
    ct_map(nc.ct_kind, GPHS_X(W::xphi(nc, i) + 1), cb0, cb1, ct, st, fct);
    factor *= fct;
    // the root's parent boost is the identity (phs_cpu.cpp:568-573): M = L1; after a branch
    // with two final-state daughters the walk pops a saved matrix, so only the columns the
    // daughters' momenta use are formed.  A branch with a friend takes its polar axis from
    // an incoming beam (phs.cu:853-910).
    if (FR && nc.friend_sign != 0) compose_friend(M, g, bg, ct, st, cp, sp, (double)nc.friend_sign);
    else if (i == 0) set_l1(M, g, bg, ct, st, cp, sp);
    else if (leaf0 && leaf1) compose_cols03(M, g, bg, ct, st, cp, sp);
    else compose(M, g, bg, ct, st, cp, sp);

    // final-state daughters: p = L (E, 0, 0, +-p) (phs_cpu.cpp:286-291), then the lab boost
    // of the final leaves (phs.cu:1286-1301); E = sqrt(0 + p^2) is |p| exactly
#pragma unroll
    for (int c = 0; c < 2; c++) {
      if (c == 0 ? leaf0 : leaf1) {
        const double pc = c == 0 ? pk : -pk;
        const double E = nc.k[c].msq_leaf == 0 ? fabs(pc) : math::sqrt_phys(nc.k[c].msq_leaf + pc * pc);
        // rows of L (E, 0, 0, pc), the second product fused
        double q0 = __builtin_fma(M[3], pc, M[0] * E);
        const double q1 = __builtin_fma(M[7], pc, M[4] * E);
        const double q2 = __builtin_fma(M[11], pc, M[8] * E);
        double q3 = __builtin_fma(M[15], pc, M[12] * E);
        if (eb.bg_lab != 0) {
          const double b0 = eb.g_lab * q0 + eb.bg_lab * q3;
          const double b3 = eb.bg_lab * q0 + eb.g_lab * q3;
          q0 = b0; q3 = b3;
        }
        double* o = pl + 4 * nc.k[c].out * S;
        o[0] = q0; o[S] = q1; o[2 * S] = q2; o[3 * S] = q3;
      }
    }

    // next branch: a daughter (matrix stays), or the deferred one from the save bank
    const int push = W::push(nc, i), desc = W::desc(nc, i);
    const double msq_push = push == 0 ? msq0 : msq1;
    const double m_push = push == 0 ? m1 : m2;
    const double p_push = push == 0 ? pk : -pk;
    if (push >= 0) { bank.push(depth, M, L1Par{g, bg, ct, st, cp, sp}, msq_push, m_push, p_push); depth++; }
    if (desc >= 0) {
      cur_msq = desc == 0 ? msq0 : msq1;
      cur_m = desc == 0 ? m1 : m2;
      cur_p = desc == 0 ? pk : -pk;
    } else if (depth > 0) {
      depth--;
      bank.pop(depth, M, cur_msq, cur_m, cur_p);
    }
  };
  if constexpr (SID >= 0) {
#pragma unroll
    for (int i = 0; i < Walk<SID>::n_nodes(cd); ++i) branch(i);
  } else {
    for (int i = 0; i < cd.n_nodes; ++i) branch(i);
  }
  factor_out = factor;
  ok_out = ok;
#undef GPHS_X
}

// Momenta of `cnt` events from ps[(4 j + mu) * S + event] to dst = [cnt][n_out][4], as
// streaming (non-temporal) stores: nothing here reads them again, and keeping them out of the
// caches measured 2-3% faster on the batched mix.  With a
// compile-time workgroup size BLK and n_out NOUT of 2 or 4 (16-byte words per event W = 2 NOUT
// dividing BLK), the word thread t stores in round k is t + BLK k: event t / W + k BLK / W,
// component pair t % W -- one address per thread, and every round's LDS and global offsets are
// constants.
template <int NOUT, int BLK>
__device__ __forceinline__ void store_momenta(const double* __restrict__ ps, int S, int cnt, const ChannelDesc& cd,
                                              bool aligned, double* __restrict__ dst) {
  if constexpr (BLK > 0 && (NOUT == 2 || NOUT == 4)) {
    constexpr int W = 2 * NOUT, SS = BLK + 1;
    if (aligned) {
      const int t = (int)threadIdx.x;
      const double* src = ps + 2 * (t % W) * SS + t / W;
      gdouble2* d2 = (gdouble2*)dst + t;
      const int nw = cnt * W;
#pragma unroll
      for (int k = 0; k < W; k++) {
        if (t + k * BLK < nw) {
          dvec2 v;
          v.x = src[k * (BLK / W)];
          v.y = src[SS + k * (BLK / W)];
          __builtin_nontemporal_store(v, d2 + k * BLK);
        }
      }
      return;
    }
  }
  const int w4 = 4 * cd.n_out;
  const int n = cnt * w4;
  if (aligned) {
    // 16-byte words; 4 n_out is even, so both doubles of a word belong to one event
    gdouble2* d2 = (gdouble2*)dst;
    for (int w = (int)threadIdx.x; w < n / 2; w += (int)blockDim.x) {
      const int i = 2 * w;
      const int ev = (int)__umulhi((unsigned)i, cd.div_p);
      const int c = i - ev * w4;
      dvec2 v;
      v.x = ps[c * S + ev];
      v.y = ps[(c + 1) * S + ev];
      __builtin_nontemporal_store(v, d2 + w);
    }
  } else {
    gdouble* d1 = as_global(dst);
    for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) {
      const int ev = (int)__umulhi((unsigned)i, cd.div_p);
      __builtin_nontemporal_store(ps[(i - ev * w4) * S + ev], d1 + i);
    }
  }
}

// WPE > 0 asks the register allocator for at least WPE waves per SIMD (fewer VGPRs, spills
// allowed); the autotuner measures which occupancy target wins.
//
// LDS per workgroup (S = blockDim.x + 1 doubles per row): momenta 4 nout_max rows, the
// randoms nx_max rows (XN = kXLds), then the save banks (BANK = BK_LDS).  A tile of `count`
// events is done in sub-tiles of blockDim.x events.
//
// BLK > 0: the workgroup size is that compile-time constant (shape kernels), so the LDS offsets
// of the momentum table and of the store phase are immediates.
//
// ONE: the launch has one workgroup per tile and tiles of at most blockDim.x events, so each
// loop below makes one pass; both end after it, so that they are no loops to the compiler,
// which would otherwise keep constants hoisted out of them in vector registers (the kernels of
// 2-4 final-state particles: 118 instead of 159 VGPRs, four waves per SIMD instead of three).
template <int BANK, int WPE, int XN, bool FR, bool PEV, int SID, int BLK, bool ONE = false>
__global__ __launch_bounds__(BLK > 0 ? BLK : 256) __attribute__((amdgpu_waves_per_eu(WPE > 0 ? WPE : 1)))
void forward_kernel(const TileDesc* __restrict__ tiles, int t_begin, int t_end, const SegDesc* __restrict__ segs,
                    const ChannelDesc* __restrict__ chans, const NodeCmd* __restrict__ nodes, int nx_max,
                    int nout_max) {
  const int B = BLK > 0 ? BLK : (int)blockDim.x;
  const int S = B + 1;
  double* ps = lds;
  double* xs = ps + 4 * nout_max * S;
  double* bank_base = xs + (XN == kXLds ? nx_max * S : 0);
  const int lane = (int)threadIdx.x;
  constexpr int kNout = SID >= 0 ? kShapes[SID < 0 ? 0 : SID].n_out : 0;
  for (int t = t_begin + (int)blockIdx.x; t < t_end; t += (int)gridDim.x) {
    const TileDesc td = tiles[t];
    const SegDesc sd = segs[td.seg];
    const ChannelDesc cd = chans[sd.channel];
    const NodeCmd* prog = nodes + cd.node0;
    // lanes past the end of a sub-tile repeat its last event and store nothing
    auto event_of = [&](int base) -> int64_t {
      const int cnt = min(B, td.count - base);
      return td.e0 + base + (lane < cnt ? lane : cnt - 1);
    };
    for (int base = 0; base < td.count; base += B) {
      const int cnt = min(B, td.count - base);
      const int64_t e0 = td.e0 + base;
      const bool valid = lane < cnt;
      double factor;
      bool ok;
      __syncthreads();   // the previous sub-tile's momentum table has been read
      const int64_t e = event_of(base);
      // (loading the next sub-tile's randoms during this one measured slower: the extra
      // registers cost a wave per SIMD)
      double xv[XN > 0 ? XN : 1];
      load_randoms<XN, SID>(sd.x + e * cd.n_x, cd.n_x, xv, xs + lane, S);
      EvBeams eb;
      if constexpr (PEV) eb = event_beams(cd, as_global(sd.e1)[e], as_global(sd.e2)[e]);
      else eb = fixed_beams(cd);
      // the msq output is wave-uniform; lanes past the end write their (repeated) last event's
      // values again, identical ones
      const bool want_msq = sd.msq != nullptr;
      gdouble* msq_row = want_msq ? as_global(sd.msq) + e * ((1 << cd.n_out) - 1) : nullptr;
      forward_event<BANK, XN, FR, PEV, SID>(cd, eb, prog, xv, ps + lane, xs + lane, S, bank_base, want_msq, msq_row,
                                            factor, ok);
      if (valid) {
        __builtin_nontemporal_store(factor, as_global(sd.factor) + e0 + lane);
        __builtin_nontemporal_store(eb.volume, as_global(sd.volume) + e0 + lane);
        as_global(sd.ok)[e0 + lane] = ok ? 1 : 0;
      }
      __syncthreads();
      store_momenta<kNout, BLK>(ps, S, cnt, cd, sd.align16 & 2, sd.p + e0 * (int64_t)cd.n_out * 4);
      if constexpr (ONE) break;
    }
    if constexpr (ONE) break;
  }
}

using FwdKernel = void (*)(const TileDesc*, int, int, const SegDesc*, const ChannelDesc*, const NodeCmd*, int, int);

// Kernels specialised for one tree shape: randoms in the shape's n_x registers, the root's
// boost parameters as the save bank when only the root defers a daughter, one level of
// registers when another branch does, fixed beams, no friends.
template <int SID, int WPE, int BLK, bool ONE = false>
FwdKernel shape_kernel() {
  constexpr const ShapeProg& sh = kShapes[SID];
  constexpr int bank = sh.banks == 0 ? BK_NONE : (sh.root_only ? BK_ROOT : BK_VGPR);
  return forward_kernel<bank, WPE, sh.n_x, false, false, SID, BLK, ONE>;
}

// The shape kernels of shapes [lo, hi) (one translation unit each): the variant for register
// budget `wpe` and workgroup size `block`, or null for a shape outside the range.
// one: the launch has one workgroup per tile and tiles of at most `block` events (the
// one-pass variants of the trees of up to four final-state particles)
FwdKernel pick_shape_0_8(int sid, int wpe, int block, bool one);
// per-event beam energies: shapes [0, kPevShapes), the trees of up to four final-state particles
FwdKernel pick_shape_pev_0_8(int sid, int wpe, int block, bool one);
FwdKernel pick_shape_8_22(int sid, int wpe, int block, bool one);
FwdKernel pick_shape_22_36(int sid, int wpe, int block, bool one);
FwdKernel pick_shape_36_50(int sid, int wpe, int block, bool one);
FwdKernel pick_shape_50_64(int sid, int wpe, int block, bool one);

}  // namespace fwd

}  // namespace gphs
