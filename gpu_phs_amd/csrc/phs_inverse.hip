// Inverse mapping kernel: for every event and every channel of its process, the randoms and
// the phase-space factor under which that channel would have produced the event's momenta
// (what Whizard's multi-channel weight needs; reference phs.cu:626-675, 924-1284, driven per
// channel by phs.cu:1499-1554 with 2 + 2(n-1) dependent launches per channel).
//
// One lane = one event, one launch for all channels: a workgroup takes a tile of events of
// one generating channel, stages their momenta (boosted back to the centre-of-mass frame) in
// an LDS column per lane, and walks every channel's program in turn (wave-uniform: all lanes
// invert into the same channel at the same time, reading its NodeCmd records with scalar
// loads).  The generating channel's slot gets the forward pass's own factor and randoms, as
// the reference's _move_factors / _move_x do (phs.cu:1495-1497).
//
// The steps per branch follow the GPU reference with its defects corrected (the oracle's
// inverse lists them): the msq of both daughters from sums of final-state momenta (added as
// the reference adds them, daughter by daughter), x and the
// map factor of each internal daughter from its msq, lambda and the decay momentum, then the
// angles of the first daughter in B_z(-bg) G, where G = L_parent^-1 is carried down the tree
// (a branch with a friend measures them from the friend beam), x_phi = atan2 / 2 pi, x of
// cos theta from its map, and G <- R_y(-theta) R_z(-phi) B_z(-bg) G for the daughters.
// Deferred subtrees wait in an LDS save bank (G and the subtree's msq, m, p_decay; the root's
// four angle values instead of G where only roots defer).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "phs_device.h"
#include "phs_kernels.h"
#include "phs_walk.h"

namespace gphs {

namespace {

constexpr int kInvBankWords = 19;

// The lane's column of the sub-tile's randoms table x[k][event] (rows of `stride` doubles).
struct XColumn {
  double* p;
  int stride;
  __device__ __forceinline__ double& operator[](int k) const { return p[k * stride]; }
};

// With RB (every deferred subtree is a daughter of the root, and no root has a friend) a bank
// entry is the root's cos theta, sin theta, cos phi, sin phi and the subtree's msq, m and
// p_decay: the root's G is its rotation applied to the identity (the root's bg is 0), so the
// pop re-forms it bit for bit from the four numbers.
constexpr int kInvRootBankWords = 7;

extern __shared__ double inv_lds[];

// x and the map factor of an internal daughter from its msq (mappings_gpu.hpp:15-172; STEP_H
// as phs_cpu.cpp:95-101, whose root of the quadratic is the right one).
__device__ __forceinline__ void x_from_msq(const ChildCmd& c, double msq, double s, double& x, double& f) {
  switch (c.kind) {
    case MK_LINEAR:
      x = (msq - c.msq_min) / (c.msq_max - c.msq_min);
      f = c.f_lin;
      break;
    case MK_BREIT_WIGNER: {
      const double z = (msq - c.m * c.m) / (c.m * c.w);
      x = (atan(z) - c.a0) / (c.a1 - c.a0);
      f = c.a2 * (1 + z * z);
      break;
    }
    case MK_EXP: {
      const double z = msq - c.msq_min + c.a0;
      x = math::log_bounded(z / c.a0) / c.a1;
      f = c.a2 * z;
      break;
    }
    case MK_TU: {
      const bool lo = msq < (c.msq_max + c.msq_min) / 2;
      const double z = lo ? msq - c.msq_min + c.a0 : c.msq_max - msq + c.a0;
      const double l = math::log_bounded(z / c.a0) / c.a1;
      x = lo ? l : 1 - l;
      f = c.a2 * z;
      break;
    }
    case MK_STEP_H: {
      const double z = (msq - c.msq_min) / (c.msq_max - c.msq_min);
      const double t = c.a1 / (c.a0 * c.a2);
      const double u = c.a0 + z / c.a2, v = c.a0 - z / c.a2;
      x = ((u + t) - math::sqrt_phys(v * v + 2 * t * u + t * t)) / 2;
      f = (c.a1 / (c.a0 - x) / (c.a0 - x) + c.a2) * (c.msq_max - c.msq_min) / s;
      break;
    }
    default: {  // MK_STEP_E
      const double z = (msq - c.msq_min) / (c.msq_max - c.msq_min);
      const double t = 1 + c.a1 * math::exp_bounded(z / c.a2);
      x = (z - c.a2 * math::log_bounded(t / (1 + c.a1))) / c.a0;
      f = c.a0 * t * (c.msq_max - c.msq_min) / s;
      break;
    }
  }
}

// x and factor of cos theta (mappings_gpu.hpp:181-219).
__device__ __forceinline__ void x_from_ct(int kind, double ct, double b0, double b1, double& x, double& f) {
  if (kind == CK_FLAT) {
    x = (ct + 1) / 2;
    f = 1;
  } else {
    const bool neg = ct < 0;
    const double t = neg ? ct + b0 + 1 : -ct + b0 + 1;
    const double l = math::log_bounded(t / b0) / (2 * b1);
    x = neg ? l : 1 - l;
    f = t * b1;
  }
}

// Sum of the staged final-state momenta in `leaves` (lane column pl, stride S).
__device__ __forceinline__ void leaf_sum(const double* __restrict__ pl, int S, int n_out, uint32_t leaves,
                                         double (&q)[4]) {
  q[0] = q[1] = q[2] = q[3] = 0;
  for (int j = 0; j < n_out; j++) {
    if ((leaves >> j) & 1u) {
      const double* r = pl + 4 * j * S;
      q[0] = q[0] + r[0]; q[1] = q[1] + r[S]; q[2] = q[2] + r[2 * S]; q[3] = q[3] + r[3 * S];
    }
  }
}

// A daughter's four-momentum: its sum program (ChildCmd::sum_code) over the staged final-state
// momenta, which adds them in the reference's association, daughter sums first; the index
// order of `leaves` where the subtree has no program.
__device__ __forceinline__ void branch_sum(const double* __restrict__ pl, int S, int n_out, const ChildCmd& c,
                                           double (&q)[4]) {
  if (c.sum_len <= 0) {
    leaf_sum(pl, S, n_out, c.leaves, q);
    return;
  }
  double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
  bool second = false;
  uint64_t code = c.sum_code;
  for (int t = 0; t < c.sum_len; t++, code >>= 4) {
    const int op = (int)(code & 15u);
    if (op == kSumPush) {
      second = true;
      b[0] = b[1] = b[2] = b[3] = 0;
    } else if (op == kSumMerge) {
      a[0] = a[0] + b[0]; a[1] = a[1] + b[1]; a[2] = a[2] + b[2]; a[3] = a[3] + b[3];
      second = false;
    } else {
      const double* r = pl + 4 * op * S;
      if (second) {
        b[0] = b[0] + r[0]; b[1] = b[1] + r[S]; b[2] = b[2] + r[2 * S]; b[3] = b[3] + r[3 * S];
      } else {
        a[0] = a[0] + r[0]; a[1] = a[1] + r[S]; a[2] = a[2] + r[2 * S]; a[3] = a[3] + r[3 * S];
      }
    }
  }
  q[0] = a[0]; q[1] = a[1]; q[2] = a[2]; q[3] = a[3];
}

// G <- R_y(-theta) R_z(-phi) G (rows 1..3)
__device__ __forceinline__ void rotate_back(double (&G)[16], double ct, double st, double cp, double sp) {
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const double w1 = G[4 + j], w2 = G[8 + j], w3 = G[12 + j];
    const double t1 = cp * w1 + sp * w2;
    G[4 + j] = ct * t1 - st * w3;
    G[8 + j] = -sp * w1 + cp * w2;
    G[12 + j] = st * t1 + ct * w3;
  }
}

// One branch's decay from its own msq and mass and its daughters' msq (phs.cu:658-664):
// p_decay, the factor sqrt(lambda) / msq (0 where lambda <= 0, phs_cpu.cpp:275-279), and the
// boost bg = p_parent / m.  The quotients are IEEE divisions, as in the reference: they feed
// the frames of the daughters and so the cosines that the peaked maps take near +-1, where an
// ulp of cos theta is amplified by up to sqrt(s) / m^2 in the map's factor.
struct Decay {
  double pk, lam, bg;
};
__device__ __forceinline__ Decay decay(double cur_msq, double cur_m, double cur_p, double msq0, double msq1) {
  const double lda = (cur_msq - msq0 - msq1) * (cur_msq - msq0 - msq1) - 4 * msq0 * msq1;
  const bool pos = lda > 0;
  const double sl = pos ? math::sqrt_phys(lda) : 0.0;
  Decay d;
  d.pk = sl / (2 * cur_m);
  d.lam = pos ? sl / cur_msq : 0.0;
  d.bg = cur_m > 0 ? cur_p / cur_m : 0.0;
  return d;
}

// Polar and azimuthal angles of the direction v (mappings_gpu.hpp:196-219 inverted): the
// cosines and sines of both angles (IEEE divisions, as above) and x_phi = atan2(v1, v0) / 2 pi
// in [0, 1], formed in turns by math::atan2_turns.
struct Dir {
  double xphi, ct, st, cp, sp;
};
__device__ __forceinline__ Dir direction(double v0, double v1, double v2) {
  const double rho = math::sqrt_phys(v0 * v0 + v1 * v1);
  const double vv = math::sqrt_phys(v0 * v0 + v1 * v1 + v2 * v2);
  Dir d;
  d.xphi = math::atan2_turns(v1, v0);
  d.cp = rho > 0 ? v0 / rho : 1.0;
  d.sp = rho > 0 ? v1 / rho : 0.0;
  d.ct = v2 / vv;
  d.st = rho / vv;
  return d;
}

// ---- shared roots --------------------------------------------------------------------------
//
// The root of every channel sits in the centre-of-mass frame (G = 1, p_decay = 0, so its boost
// is the identity and v = the first daughter's momentum, bit for bit), and what it computes from
// the event -- both daughters' msq, lambda, p_decay and the first daughter's angles -- depends
// only on which final-state particles each daughter holds and in which association they are
// added.  Channels whose roots agree in both (the same root split up to the orientation of the
// subtrees below it) share one evaluation: the host orders the channels so that they follow one
// another, and the kernel evaluates the root when a new group starts.  Each channel then only
// applies its own maps to the shared numbers (x and factor of its internal daughters' msq, x
// and factor of cos theta) and walks the rest of its tree from the root's frame, which is the
// rotation R_y(-theta) R_z(-phi) of the identity, re-formed from four numbers.  That frame and
// the group's msq, m and p_decay also stand in for a subtree the root defers, so the save bank
// holds only subtrees deferred below the root.

struct RootGeo {
  double msq[2];          // msq of the daughters, from the sums of their final-state momenta
  double m[2];            // sqrt(msq) of an internal daughter (unused for a final-state one)
  double lam;             // sqrt(lambda) / s, 0 where lambda <= 0 (phs_cpu.cpp:275-279)
  double pk;              // p_decay = sqrt(lambda) / (2 sqrt(s))
  double xphi;            // phi / 2 pi of daughter 0
  double ct, st, cp, sp;  // cos and sin of its polar and azimuthal angles
};

__device__ __forceinline__ RootGeo root_geometry(const NodeCmd& nc, const EvBeams& eb,
                                                 const double* __restrict__ pl, int S, int n_out) {
  RootGeo r;
  double q0[4], q1[4];
  branch_sum(pl, S, n_out, nc.k[0], q0);
  branch_sum(pl, S, n_out, nc.k[1], q1);
  r.msq[0] = q0[0] * q0[0] - q0[1] * q0[1] - q0[2] * q0[2] - q0[3] * q0[3];
  r.msq[1] = q1[0] * q1[0] - q1[1] * q1[1] - q1[2] * q1[2] - q1[3] * q1[3];
#pragma unroll
  for (int d = 0; d < 2; d++) r.m[d] = nc.k[d].leaf ? 0.0 : math::sqrt_phys(r.msq[d]);
  const Decay dk = decay(eb.s, eb.sqrts, 0.0, r.msq[0], r.msq[1]);
  r.pk = dk.pk;
  r.lam = dk.lam;
  // G = 1 and bg = 0 at the root: v is daughter 0's momentum
  const Dir dr = direction(q0[1], q0[2], q0[3]);
  r.xphi = dr.xphi;
  r.ct = dr.ct; r.st = dr.st; r.cp = dr.cp; r.sp = dr.sp;
  return r;
}

// Momentum sum of the subtree below visit index V of shape SID, in the reference's association
// (each branch the sum of its two daughters, phs.cu:1262-1275; the sum programs of branch_sum
// add the same pairs, in an order that differs only by commutation).
template <int SID, int V>
__device__ __forceinline__ void subtree_sum(const NodeCmd* __restrict__ prog, const double* __restrict__ pl, int S,
                                            double (&q)[4]) {
  constexpr ShapeNode N = kShapes[SID].node[V];
  double a[4], b[4];
  if constexpr (N.leaf0) {
    const double* r = pl + 4 * prog[V].k[0].out * S;
    a[0] = r[0]; a[1] = r[S]; a[2] = r[2 * S]; a[3] = r[3 * S];
  } else {
    subtree_sum<SID, (N.node0 < 0 ? 0 : N.node0)>(prog, pl, S, a);
  }
  if constexpr (N.leaf1) {
    const double* r = pl + 4 * prog[V].k[1].out * S;
    b[0] = r[0]; b[1] = r[S]; b[2] = r[2 * S]; b[3] = r[3 * S];
  } else {
    subtree_sum<SID, (N.node1 < 0 ? 0 : N.node1)>(prog, pl, S, b);
  }
  q[0] = a[0] + b[0]; q[1] = a[1] + b[1]; q[2] = a[2] + b[2]; q[3] = a[3] + b[3];
}

// Daughter c of branch i: a final-state momentum, or its subtree's sum.
template <int SID, int I, int C>
__device__ __forceinline__ void daughter_sum(const NodeCmd* __restrict__ prog, const double* __restrict__ pl, int S,
                                             double (&q)[4]) {
  constexpr ShapeNode N = kShapes[SID].node[I];
  if constexpr (C == 0 ? N.leaf0 : N.leaf1) {
    const double* r = pl + 4 * prog[I].k[C].out * S;
    q[0] = r[0]; q[1] = r[S]; q[2] = r[2 * S]; q[3] = r[3 * S];
  } else {
    constexpr int child = C == 0 ? N.node0 : N.node1;
    subtree_sum<SID, (child < 0 ? 0 : child)>(prog, pl, S, q);
  }
}

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}

// The inverse of one event into channel `cd` (program `prog`) from its root group's geometry
// `g`: the randoms go to xo[0..n_x), returns the factor.  SID >= 0: the channel's program has
// shape SID of the compile-time table and every branch's structure is a constant; SID < 0: it
// is read from the program.  No branch has a friend.
template <int SID, bool PEV, class XO>
__device__ __forceinline__ double inverse_channel(const ChannelDesc& cd, const EvBeams& eb,
                                                  const NodeCmd* __restrict__ prog, const RootGeo& g,
                                                  const double* __restrict__ pl, int S, double* bank, int B,
                                                  XO xo) {
  using W = Walk<SID>;
  const double s = eb.s;
  double factor = 1;
  // the root: this channel's maps of the shared numbers
  {
    const NodeCmd& nc = prog[0];
#pragma unroll
    for (int d = 0; d < 2; d++) {
      if (!W::leaf(nc, 0, d)) {
        double x, f;
        if constexpr (PEV) x_from_msq(child_at(nc.k[d], eb.sqrts, s, cd.m_tot), g.msq[d], s, x, f);
        else x_from_msq(nc.k[d], g.msq[d], s, x, f);
        xo[W::xi(nc, 0, d)] = x;
        factor *= f;
      }
    }
    factor *= g.lam;
    xo[W::xphi(nc, 0)] = g.xphi;
    double xc, fct, m_max, cb0 = nc.b0, cb1 = nc.b1;
    if constexpr (PEV) node_at(nc, true, eb, cd.m_tot, m_max, cb0, cb1);
    x_from_ct(nc.ct_kind, g.ct, cb0, cb1, xc, fct);
    xo[W::xphi(nc, 0) + 1] = xc;
    factor *= fct;
  }
  if (W::n_nodes(cd) == 1) return factor;

  // the root's frame, and the daughter the walk descends into
  double G[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  rotate_back(G, g.ct, g.st, g.cp, g.sp);
  const int rdesc = W::desc(prog[0], 0);
  double cur_msq = g.msq[rdesc], cur_m = g.m[rdesc], cur_p = rdesc == 0 ? g.pk : -g.pk;
  bool root_pending = W::push(prog[0], 0) >= 0;
  int depth = 0;   // save-bank entries of subtrees deferred below the root

  auto branch = [&](auto ic) {
    const int i = ic;
    const NodeCmd& nc = prog[i];
    double q0[4], q1[4];
    if constexpr (SID >= 0) {
      daughter_sum<(SID < 0 ? 0 : SID), decltype(ic)::value, 0>(prog, pl, S, q0);
      daughter_sum<(SID < 0 ? 0 : SID), decltype(ic)::value, 1>(prog, pl, S, q1);
    } else {
      branch_sum(pl, S, cd.n_out, nc.k[0], q0);
      branch_sum(pl, S, cd.n_out, nc.k[1], q1);
    }
    const double msq0 = q0[0] * q0[0] - q0[1] * q0[1] - q0[2] * q0[2] - q0[3] * q0[3];
    const double msq1 = q1[0] * q1[0] - q1[1] * q1[1] - q1[2] * q1[2] - q1[3] * q1[3];
#pragma unroll
    for (int d = 0; d < 2; d++) {
      if (!W::leaf(nc, i, d)) {
        double x, f;
        if constexpr (PEV) x_from_msq(child_at(nc.k[d], eb.sqrts, s, cd.m_tot), d == 0 ? msq0 : msq1, s, x, f);
        else x_from_msq(nc.k[d], d == 0 ? msq0 : msq1, s, x, f);
        xo[W::xi(nc, i, d)] = x;
        factor *= f;
      }
    }
    const Decay dk = decay(cur_msq, cur_m, cur_p, msq0, msq1);
    const double pk = dk.pk, bg = dk.bg;
    factor *= dk.lam;

    const double gm = math::sqrt_phys(1 + bg * bg);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const double r0 = G[j], r3 = G[12 + j];
      G[j] = gm * r0 - bg * r3;
      G[12 + j] = -bg * r0 + gm * r3;
    }
    const double v0 = G[4] * q0[0] + G[5] * q0[1] + G[6] * q0[2] + G[7] * q0[3];
    const double v1 = G[8] * q0[0] + G[9] * q0[1] + G[10] * q0[2] + G[11] * q0[3];
    const double v2 = G[12] * q0[0] + G[13] * q0[1] + G[14] * q0[2] + G[15] * q0[3];
    const Dir dr = direction(v0, v1, v2);
    const double ct = dr.ct, st = dr.st, cp = dr.cp, sp = dr.sp;
    xo[W::xphi(nc, i)] = dr.xphi;
    double xc, fct, m_max, cb0 = nc.b0, cb1 = nc.b1;
    if constexpr (PEV) node_at(nc, false, eb, cd.m_tot, m_max, cb0, cb1);
    x_from_ct(nc.ct_kind, ct, cb0, cb1, xc, fct);
    xo[W::xphi(nc, i) + 1] = xc;
    factor *= fct;
    rotate_back(G, ct, st, cp, sp);

    const int push = W::push(nc, i), desc = W::desc(nc, i);
    if (push >= 0) {
      const double mp = push == 0 ? msq0 : msq1;
      double* b = bank + kInvBankWords * depth * B;
#pragma unroll
      for (int j = 0; j < 16; j++) b[j * B] = G[j];
      b[16 * B] = mp;
      b[17 * B] = math::sqrt_phys(mp);
      b[18 * B] = push == 0 ? pk : -pk;
      depth++;
    }
    if (desc >= 0) {
      cur_msq = desc == 0 ? msq0 : msq1;
      cur_m = math::sqrt_phys(cur_msq);
      cur_p = desc == 0 ? pk : -pk;
    } else if (depth > 0) {
      depth--;
      const double* b = bank + kInvBankWords * depth * B;
#pragma unroll
      for (int j = 0; j < 16; j++) G[j] = b[j * B];
      cur_msq = b[16 * B];
      cur_m = b[17 * B];
      cur_p = b[18 * B];
    } else if (root_pending) {
      // the subtree the root deferred: the root's frame again, re-formed bit for bit
      root_pending = false;
      const int rp = 1 - rdesc;
#pragma unroll
      for (int j = 0; j < 16; j++) G[j] = (j % 5 == 0) ? 1.0 : 0.0;
      rotate_back(G, g.ct, g.st, g.cp, g.sp);
      cur_msq = g.msq[rp];
      cur_m = g.m[rp];
      cur_p = rp == 0 ? g.pk : -g.pk;
    }
  };
  if constexpr (SID >= 0) {
    static_for<1, kShapes[SID < 0 ? 0 : SID].n_nodes>(branch);
  } else {
    for (int i = 1; i < cd.n_nodes; ++i) branch(i);
  }
  return factor;
}

// Inverse of one event into channel `cd` (program `prog`): the randoms go to xo[0..n_x),
// returns the factor.
template <bool FR, bool PEV, bool RB>
__device__ __forceinline__ double inverse_event(const ChannelDesc& cd, const EvBeams& eb,
                                                const NodeCmd* __restrict__ prog, const double* __restrict__ pl,
                                                int S, double* bank, int B, gdouble* __restrict__ xo) {
  const double s = eb.s;
  double G[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  double cur_msq = s, cur_m = eb.sqrts, cur_p = 0, factor = 1;
  int depth = 0;
  for (int i = 0; i < cd.n_nodes; ++i) {
    const NodeCmd& nc = prog[i];
    double q0[4], q1[4];
    branch_sum(pl, S, cd.n_out, nc.k[0], q0);
    branch_sum(pl, S, cd.n_out, nc.k[1], q1);
    const double msq0 = q0[0] * q0[0] - q0[1] * q0[1] - q0[2] * q0[2] - q0[3] * q0[3];
    const double msq1 = q1[0] * q1[0] - q1[1] * q1[1] - q1[2] * q1[2] - q1[3] * q1[3];
    // x and factor of the internal daughters' msq maps
#pragma unroll
    for (int d = 0; d < 2; d++) {
      if (!nc.k[d].leaf) {
        double x, f;
        if constexpr (PEV) x_from_msq(child_at(nc.k[d], eb.sqrts, s, cd.m_tot), d == 0 ? msq0 : msq1, s, x, f);
        else x_from_msq(nc.k[d], d == 0 ? msq0 : msq1, s, x, f);
        xo[nc.k[d].xi] = x;
        factor *= f;
      }
    }
    // this branch's decay
    const Decay dk = decay(cur_msq, cur_m, cur_p, msq0, msq1);
    const double pk = dk.pk, bg = dk.bg;
    factor *= dk.lam;

    // angles of the first daughter in W = B_z(-bg) G
    const double g = math::sqrt_phys(1 + bg * bg);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const double r0 = G[j], r3 = G[12 + j];
      G[j] = g * r0 - bg * r3;
      G[12 + j] = -bg * r0 + g * r3;
    }
    double v0 = G[4] * q0[0] + G[5] * q0[1] + G[6] * q0[2] + G[7] * q0[3];
    double v1 = G[8] * q0[0] + G[9] * q0[1] + G[10] * q0[2] + G[11] * q0[3];
    double v2 = G[12] * q0[0] + G[13] * q0[1] + G[14] * q0[2] + G[15] * q0[3];
    if (FR && nc.friend_sign != 0) {
      // the beam k = (1, 0, 0, f) in the same frame fixes R (R e_z = n / |n|); v <- R^T v and
      // the rows of W likewise, so that G below is R_y(-theta) R_z(-phi) R^T W
      const double fs = (double)nc.friend_sign;
      const double n0 = G[4] + fs * G[7], n1 = G[8] + fs * G[11], n2 = G[12] + fs * G[15];
      const double nn = math::sqrt_phys(n0 * n0 + n1 * n1 + n2 * n2);
      const double a0 = n0 / nn, a1 = n1 / nn, c = n2 / nn;
      const double sa = math::sqrt_phys(a0 * a0 + a1 * a1);
      const double kx = sa > 0 ? -a1 / sa : 1.0, ky = sa > 0 ? a0 / sa : 0.0, h = 1 - c;
      const double r00 = c + h * kx * kx, r01 = h * kx * ky, r11 = c + h * ky * ky;
      // R^T rows: (r00, r01, -a0), (r01, r11, -a1), (a0, a1, c)
      const double u0 = r00 * v0 + r01 * v1 + -a0 * v2;
      const double u1 = r01 * v0 + r11 * v1 + -a1 * v2;
      const double u2 = a0 * v0 + a1 * v1 + c * v2;
      v0 = u0; v1 = u1; v2 = u2;
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const double w1 = G[4 + j], w2 = G[8 + j], w3 = G[12 + j];
        G[4 + j] = r00 * w1 + r01 * w2 + -a0 * w3;
        G[8 + j] = r01 * w1 + r11 * w2 + -a1 * w3;
        G[12 + j] = a0 * w1 + a1 * w2 + c * w3;
      }
    }
    const Dir dr = direction(v0, v1, v2);
    const double ct = dr.ct, st = dr.st, cp = dr.cp, sp = dr.sp;
    xo[nc.x_phi] = dr.xphi;
    double xc, fct, m_max, cb0 = nc.b0, cb1 = nc.b1;
    if constexpr (PEV) node_at(nc, i == 0, eb, cd.m_tot, m_max, cb0, cb1);
    x_from_ct(nc.ct_kind, ct, cb0, cb1, xc, fct);
    xo[nc.x_phi + 1] = xc;
    factor *= fct;
    rotate_back(G, ct, st, cp, sp);

    // next branch: a daughter (G stays), or the deferred one from the save bank
    if (nc.push >= 0) {
      const double mp = nc.push == 0 ? msq0 : msq1;
      if constexpr (RB) {   // only the root pushes
        double* b = bank;
        b[0] = ct; b[B] = st; b[2 * B] = cp; b[3 * B] = sp;
        b[4 * B] = mp;
        b[5 * B] = math::sqrt_phys(mp);
        b[6 * B] = nc.push == 0 ? pk : -pk;
      } else {
        double* b = bank + kInvBankWords * depth * B;
#pragma unroll
        for (int j = 0; j < 16; j++) b[j * B] = G[j];
        b[16 * B] = mp;
        b[17 * B] = math::sqrt_phys(mp);
        b[18 * B] = nc.push == 0 ? pk : -pk;
      }
      depth++;
    }
    if (nc.desc >= 0) {
      cur_msq = nc.desc == 0 ? msq0 : msq1;
      cur_m = math::sqrt_phys(cur_msq);
      cur_p = nc.desc == 0 ? pk : -pk;
    } else if (depth > 0) {
      depth--;
      if constexpr (RB) {
        const double* b = bank;
#pragma unroll
        for (int j = 0; j < 16; j++) G[j] = (j % 5 == 0) ? 1.0 : 0.0;
        rotate_back(G, b[0], b[B], b[2 * B], b[3 * B]);
        cur_msq = b[4 * B];
        cur_m = b[5 * B];
        cur_p = b[6 * B];
      } else {
        const double* b = bank + kInvBankWords * depth * B;
#pragma unroll
        for (int j = 0; j < 16; j++) G[j] = b[j * B];
        cur_msq = b[16 * B];
        cur_m = b[17 * B];
        cur_p = b[18 * B];
      }
    }
  }
  return factor;
}

// LDS per workgroup: the momenta column (4 n_out rows of blockDim.x + 1 doubles), then the
// save banks (19 n_banks rows of blockDim.x doubles).
template <bool FR, bool PEV, bool RB>
__global__ __launch_bounds__(256) void inverse_kernel(const TileDesc* __restrict__ tiles, int n_tiles,
                                                      const SegDesc* __restrict__ segs,
                                                      const InvSeg* __restrict__ isegs,
                                                      const ChannelDesc* __restrict__ chans,
                                                      const NodeCmd* __restrict__ nodes, int n_ch, int n_out,
                                                      int n_x, int64_t n_rows, double* __restrict__ factors,
                                                      double* __restrict__ x_out) {
  const int B = (int)blockDim.x;
  const int S = B + 1;
  const int lane = (int)threadIdx.x;
  double* ps = inv_lds;
  double* bank = ps + 4 * n_out * S + lane;
  double* pl = ps + lane;
  for (int t = (int)blockIdx.x; t < n_tiles; t += (int)gridDim.x) {
    const TileDesc td = tiles[t];
    const SegDesc sd = segs[td.seg];
    const InvSeg is = isegs[td.seg];
    for (int base = 0; base < td.count; base += B) {
      const int cnt = min(B, td.count - base);
      const bool valid = lane < cnt;   // lanes past the end stage the last event and store nothing
      const int64_t e = td.e0 + base + (valid ? lane : cnt - 1);
      const int64_t row = is.row0 + e;
      EvBeams eb;
      if constexpr (PEV) eb = event_beams(chans[0], as_global(sd.e1)[e], as_global(sd.e2)[e]);
      else eb = fixed_beams(chans[0]);
      // momenta back in the centre-of-mass frame (inverse of phs.cu:1286-1301)
      const gdouble* pe = as_global((const double*)sd.p) + e * 4 * n_out;
      for (int j = 0; j < n_out; j++) {
        double q0 = pe[4 * j], q1 = pe[4 * j + 1], q2 = pe[4 * j + 2], q3 = pe[4 * j + 3];
        if (eb.bg_lab != 0) {
          const double a = eb.g_lab * q0 - eb.bg_lab * q3;
          const double b = -eb.bg_lab * q0 + eb.g_lab * q3;
          q0 = a; q3 = b;
        }
        pl[(4 * j) * S] = q0; pl[(4 * j + 1) * S] = q1; pl[(4 * j + 2) * S] = q2; pl[(4 * j + 3) * S] = q3;
      }
      for (int c = 0; c < n_ch; c++) {
        gdouble* xo = as_global(x_out) + ((int64_t)c * n_rows + row) * n_x;
        if (c == is.own) {
          if (valid) {
            const gdouble* xi = as_global(sd.x) + e * n_x;
            for (int k = 0; k < n_x; k++) xo[k] = xi[k];
            as_global(factors)[row * n_ch + c] = as_global((const double*)sd.factor)[e];
          }
          continue;
        }
        const ChannelDesc cd = chans[c];
        if (valid) as_global(factors)[row * n_ch + c] = inverse_event<FR, PEV, RB>(cd, eb, nodes + cd.node0, pl, S, bank, B, xo);
      }
    }
  }
}

// Shared roots (processes without friends): the channels in the order of `steps`, a root
// evaluated where a group starts.  Channels of up to four final-state particles run the walk
// specialised for their shape; the others the walk that reads its structure from the program,
// which only the GEN instances contain (it needs some thirty more registers than the
// specialised walks).  WPE: waves per SIMD the register allocation aims at (0: the compiler's
// choice).
// LDS per workgroup: the momenta column (4 n_out rows of blockDim.x + 1 doubles), the save
// banks of subtrees deferred below the root (19 rows of blockDim.x doubles per level), then the
// randoms table of the channel being inverted (n_x rows of blockDim.x + 1 doubles).
template <bool PEV, bool GEN, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE))) void inverse_kernel_shared(const TileDesc* __restrict__ tiles, int n_tiles,
                                                             const SegDesc* __restrict__ segs,
                                                             const InvSeg* __restrict__ isegs,
                                                             const InvStep* __restrict__ steps,
                                                             const ChannelDesc* __restrict__ chans,
                                                             const NodeCmd* __restrict__ nodes, int n_ch, int n_out,
                                                             int n_x, int64_t n_rows, int rest_banks,
                                                             double* __restrict__ factors,
                                                             double* __restrict__ x_out) {
  const int B = (int)blockDim.x;
  const int S = B + 1;
  const int lane = (int)threadIdx.x;
  double* ps = inv_lds;
  double* bank = ps + 4 * n_out * S + lane;
  double* pl = ps + lane;
  // each channel's randoms of the sub-tile go through the table xs[k][event], then to x_out as
  // one contiguous run of rows (whole lines, streaming stores) instead of 8-byte pieces 8 n_x
  // bytes apart
  double* xs = ps + 4 * n_out * S + kInvBankWords * rest_banks * B;
  const XColumn xl{xs + lane, S};
  const int q_step_r = B / n_x, q_step_k = B % n_x;
  for (int t = (int)blockIdx.x; t < n_tiles; t += (int)gridDim.x) {
    const TileDesc td = tiles[t];
    const SegDesc sd = segs[td.seg];
    const InvSeg is = isegs[td.seg];
    for (int base = 0; base < td.count; base += B) {
      const int cnt = min(B, td.count - base);
      const bool valid = lane < cnt;   // lanes past the end stage the last event and store nothing
      const int64_t e = td.e0 + base + (valid ? lane : cnt - 1);
      const int64_t row = is.row0 + e;
      EvBeams eb;
      if constexpr (PEV) eb = event_beams(chans[0], as_global(sd.e1)[e], as_global(sd.e2)[e]);
      else eb = fixed_beams(chans[0]);
      const gdouble* pe = as_global((const double*)sd.p) + e * 4 * n_out;
      for (int j = 0; j < n_out; j++) {
        double q0 = pe[4 * j], q1 = pe[4 * j + 1], q2 = pe[4 * j + 2], q3 = pe[4 * j + 3];
        if (eb.bg_lab != 0) {
          const double a = eb.g_lab * q0 - eb.bg_lab * q3;
          const double b = -eb.bg_lab * q0 + eb.g_lab * q3;
          q0 = a; q3 = b;
        }
        pl[(4 * j) * S] = q0; pl[(4 * j + 1) * S] = q1; pl[(4 * j + 2) * S] = q2; pl[(4 * j + 3) * S] = q3;
      }
      RootGeo geo;
      for (int k = 0; k < n_ch; k++) {
        const InvStep st = steps[k];
        const int c = st.channel;
        const ChannelDesc cd = chans[c];
        const NodeCmd* prog = nodes + cd.node0;
        if (st.new_group) geo = root_geometry(prog[0], eb, pl, S, n_out);
        if (c == is.own) {
          if (valid) {
            const gdouble* xi = as_global(sd.x) + e * n_x;
            for (int q = 0; q < n_x; q++) xl[q] = xi[q];
            as_global(factors)[row * n_ch + c] = as_global((const double*)sd.factor)[e];
          }
        } else if (valid) {
          double f;
          switch (cd.shape) {
            case 0: f = inverse_channel<0, PEV>(cd, eb, prog, geo, pl, S, bank, B, xl); break;
            case 1: f = inverse_channel<1, PEV>(cd, eb, prog, geo, pl, S, bank, B, xl); break;
            case 2: f = inverse_channel<2, PEV>(cd, eb, prog, geo, pl, S, bank, B, xl); break;
            case 3: f = inverse_channel<3, PEV>(cd, eb, prog, geo, pl, S, bank, B, xl); break;
            case 4: f = inverse_channel<4, PEV>(cd, eb, prog, geo, pl, S, bank, B, xl); break;
            case 5: f = inverse_channel<5, PEV>(cd, eb, prog, geo, pl, S, bank, B, xl); break;
            case 6: f = inverse_channel<6, PEV>(cd, eb, prog, geo, pl, S, bank, B, xl); break;
            case 7: f = inverse_channel<7, PEV>(cd, eb, prog, geo, pl, S, bank, B, xl); break;
            default:
              if constexpr (GEN) f = inverse_channel<-1, PEV>(cd, eb, prog, geo, pl, S, bank, B, xl);
              else f = 0;   // not reached: the host runs channels of other programs with GEN
              break;
          }
          as_global(factors)[row * n_ch + c] = f;
        }
        // the sub-tile's rows of channel c, x_out[c][row0 + e0 + base ...][n_x]: element q of
        // the run is row r = q / n_x, random k = q % n_x of the table
        __syncthreads();
        gdouble* dst = as_global(x_out) + ((int64_t)c * n_rows + is.row0 + td.e0 + base) * n_x;
        int r = lane / n_x, kx = lane % n_x;
        for (int q = lane; q < cnt * n_x; q += B) {
          __builtin_nontemporal_store(xs[kx * S + r], dst + q);
          r += q_step_r;
          kx += q_step_k;
          if (kx >= n_x) { kx -= n_x; r++; }
        }
        __syncthreads();
      }
    }
  }
}

}  // namespace

size_t inverse_lds_bytes(const InverseTables& it, int block) {
  if (it.shared_roots)
    return ((size_t)4 * it.n_out * (block + 1) + (size_t)kInvBankWords * it.rest_banks * block +
            (size_t)it.n_x * (block + 1)) * sizeof(double);
  const size_t words = it.root_bank ? kInvRootBankWords : kInvBankWords;
  return ((size_t)4 * it.n_out * (block + 1) + words * it.n_banks * block) * sizeof(double);
}

hipError_t launch_inverse(const InverseTables& it, const TileDesc* tiles, int n_tiles, const SegDesc* segs,
                          int block, int grid, double* factors, double* x_out, int64_t n_rows, hipStream_t stream) {
  if (n_tiles <= 0) return hipSuccess;
  if (block < 64 || block > 256 || block % 64) return hipErrorInvalidValue;
  const size_t shmem = inverse_lds_bytes(it, block);
  if (shmem > forward_lds_limit()) return hipErrorInvalidValue;
  const int g = grid > 0 ? (grid < n_tiles ? grid : n_tiles) : n_tiles;
  if (it.shared_roots) {
    auto k = it.per_event_beams ? (it.generic ? inverse_kernel_shared<true, true, 0> : inverse_kernel_shared<true, false, 3>)
                                : (it.generic ? inverse_kernel_shared<false, true, 0> : inverse_kernel_shared<false, false, 3>);
    hipLaunchKernelGGL(k, dim3(g), dim3(block), shmem, stream, tiles, n_tiles, segs, it.isegs, it.steps, it.chans,
                       it.nodes, it.n_ch, it.n_out, it.n_x, n_rows, it.rest_banks, factors, x_out);
    return hipGetLastError();
  }
  using K = decltype(&inverse_kernel<false, false, false>);
  K k;
  if (it.root_bank)
    k = it.per_event_beams ? (it.has_friends ? inverse_kernel<true, true, true> : inverse_kernel<false, true, true>)
                           : (it.has_friends ? inverse_kernel<true, false, true> : inverse_kernel<false, false, true>);
  else
    k = it.per_event_beams ? (it.has_friends ? inverse_kernel<true, true, false> : inverse_kernel<false, true, false>)
                           : (it.has_friends ? inverse_kernel<true, false, false> : inverse_kernel<false, false, false>);
  hipLaunchKernelGGL(k, dim3(g), dim3(block), shmem, stream, tiles, n_tiles, segs, it.isegs, it.chans, it.nodes,
                     it.n_ch, it.n_out, it.n_x, n_rows, factors, x_out);
  return hipGetLastError();
}

}  // namespace gphs
