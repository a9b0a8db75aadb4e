// Reverse-mode derivatives of one event: d(momenta, factor)/d(randoms, M, Gamma).
//
// No counterpart in the reference.  The forward pass is the one of phs_forward.hip (the
// reference's phs_cpu.cpp algorithm, visit order of the compiled program); it records a
// small tape per branch: the six numbers of its boost L1 = B_z(bg) R_z(phi) R_y(theta)
// (gamma, bg, cos theta, sin theta, cos phi, sin phi) and the branch's incoming msq and
// p_decay.  A branch's parent boost is the product of the L1 of its ancestors, re-formed from
// the tape where it is needed (a deferred subtree in the forward pass, every branch in the
// reverse sweep), so no 4x4 matrix is stored per branch.  The reverse sweep visits the
// branches backwards; the adjoint of a deferred subtree waits in a stash until its parent is
// reached.  The maps and lambda are recomputed from the randoms instead of stored.
//
// Everything here is __host__ __device__ so the same code can be checked on the CPU.
#pragma once
#include <hip/hip_runtime.h>

#include "phs_device.h"
#include "phs_kernels.h"
#include "phs_mapconst.h"
#include "phs_math.h"
#include "phs_types.h"
#include "phs_walk.h"

namespace gphs {

#define GPHS_HD __host__ __device__ __forceinline__

constexpr double kTwoPi = 6.28318530717958647693;
constexpr double kPi = 3.14159265358979323846;

// msq map of an internal daughter with the intermediates its derivative needs.
struct MapEval {
  double msq, f;
  double t;     // BW: tan z; EXP/TU: exp(u a1); STEP_E: tmp; STEP_H: 1/(a0 - x)
  double aux;   // TU: sign (+1 lower half, -1 upper); STEP_E: exp(-x a0 / a2)
  bool in;      // BW: z inside (-pi/2, pi/2)
};

GPHS_HD MapEval msq_eval(const ChildCmd& c, double x, double s) {
  MapEval r{0, 0, 0, 0, true};
  switch (c.kind) {
    case MK_LINEAR:
      r.msq = (1 - x) * c.msq_min + x * c.msq_max;
      r.f = c.f_lin;
      break;
    case MK_BREIT_WIGNER: {
      double z = (1 - x) * c.a0 + x * c.a1;
      r.t = math::tan_radians(z);
      r.in = (-kPi / 2 < z) && (z < kPi / 2);
      r.msq = r.in ? c.m * (c.m + c.w * r.t) : 0.0;
      r.f = r.in ? c.a2 * (1 + r.t * r.t) : 0.0;
      break;
    }
    case MK_EXP: {
      r.t = math::exp_bounded(x * c.a1);
      double m1 = c.a0 * r.t;
      r.msq = m1 - c.a0 + c.msq_min;
      r.f = c.a2 * m1;
      break;
    }
    case MK_TU: {
      bool lo = x < 0.5;
      r.aux = lo ? 1.0 : -1.0;
      r.t = math::exp_bounded((lo ? x : (1 - x)) * c.a1);
      double m1 = c.a0 * r.t;
      r.msq = lo ? (m1 - c.a0 + c.msq_min) : (-(m1 - c.a0) + c.msq_max);
      r.f = c.a2 * m1;
      break;
    }
    case MK_STEP_H: {
      double z = c.a1 / (c.a0 - x) - c.a1 / c.a0 + c.a2 * x;
      r.t = 1 / (c.a0 - x);
      r.msq = z * c.msq_max + (1 - z) * c.msq_min;
      r.f = (c.a1 / ((c.a0 - x) * (c.a0 - x)) + c.a2) * (c.msq_max - c.msq_min) / s;
      break;
    }
    default: {  // MK_STEP_E
      r.aux = math::exp_bounded(-x * c.a0 / c.a2);
      r.t = r.aux * (1 + c.a1);
      double z = -c.a2 * math::log_bounded(r.t - c.a1);
      r.msq = z * c.msq_max + (1 - z) * c.msq_min;
      r.f = c.a0 / (1 - c.a1 / r.t) * (c.msq_max - c.msq_min) / s;
      break;
    }
  }
  return r;
}

// Adjoint of msq_eval: given gmsq, gf, d/dx, the adjoints of the map constants a[0..2] and the
// direct d/d(M, Gamma) of the map (through its own m and w, not through the constants).
GPHS_HD void msq_eval_adj_core(const ChildCmd& c, double x, double s, const MapEval& r, double gmsq, double gf,
                               double& gx, double& ga0, double& ga1, double& ga2, double& gm_direct,
                               double& gw_direct) {
#pragma clang fp contract(fast)
  ga0 = ga1 = ga2 = gm_direct = gw_direct = 0;
  gx = 0;
  switch (c.kind) {
    case MK_LINEAR:
      gx = gmsq * (c.msq_max - c.msq_min);
      break;
    case MK_BREIT_WIGNER:
      if (r.in) {
        const double sec2 = 1 + r.t * r.t;
        const double gt = gmsq * c.m * c.w + gf * c.a2 * 2 * r.t;
        const double gz = gt * sec2;
        gx = gz * (c.a1 - c.a0);
        ga0 = gz * (1 - x);
        ga1 = gz * x;
        ga2 = gf * sec2;
        gm_direct = gmsq * (2 * c.m + c.w * r.t);
        gw_direct = gmsq * c.m * r.t;
      }
      break;
    case MK_EXP: {
      const double m1 = c.a0 * r.t;
      const double gm1 = gmsq + gf * c.a2;
      ga0 = gm1 * r.t - gmsq;
      ga1 = gm1 * m1 * x;
      gx = gm1 * m1 * c.a1;
      ga2 = gf * m1;
      break;
    }
    case MK_TU: {
      const double m1 = c.a0 * r.t;
      const double sg = r.aux;                 // msq = sg * m1 - sg * a0 + const
      const double u = sg > 0 ? x : 1 - x;
      const double gm1 = sg * gmsq + gf * c.a2;
      ga0 = gm1 * r.t - sg * gmsq;
      ga1 = gm1 * m1 * u;
      gx = sg * gm1 * m1 * c.a1;
      ga2 = gf * m1;
      break;
    }
    case MK_STEP_H: {
      const double d = c.msq_max - c.msq_min, cs = d / s, rr = r.t;
      const double gz = gmsq * d;
      gx = gz * (c.a1 * rr * rr + c.a2) + gf * 2 * c.a1 * rr * rr * rr * cs;
      ga0 = gz * (-c.a1 * rr * rr + c.a1 / (c.a0 * c.a0)) - gf * 2 * c.a1 * rr * rr * rr * cs;
      ga1 = gz * (rr - 1 / c.a0) + gf * rr * rr * cs;
      ga2 = gz * x + gf * cs;
      break;
    }
    default: {  // MK_STEP_E: tmp = e (1 + a1), v = tmp - a1, z = -a2 log v, f = cs a0 tmp / v
      const double d = c.msq_max - c.msq_min, cs = d / s;
      const double tmp = r.t, e = r.aux, v = tmp - c.a1;
      const double gz = gmsq * d;
      const double gv = gz * (-c.a2 / v);
      const double gtmp = gv + gf * (-cs * c.a0 * c.a1 / (v * v));
      ga1 = -gv + gf * (cs * c.a0 * tmp / (v * v)) + gtmp * e;
      ga0 = gf * cs * tmp / v + gtmp * tmp * (-x / c.a2);
      ga2 = gz * (-math::log_bounded(v)) + gtmp * tmp * x * c.a0 / (c.a2 * c.a2);
      gx = gtmp * tmp * (-c.a0 / c.a2);
      break;
    }
  }
}

// Adjoint of msq_eval: given gmsq, gf, accumulate d/dx and d/d(M, Gamma) of the map.
GPHS_HD void msq_eval_adj(const ChildCmd& c, double x, double s, const MapEval& r, double gmsq, double gf, double& gx,
                          double& gM, double& gW) {
#pragma clang fp contract(fast)
  double ga0, ga1, ga2, gm_direct, gw_direct;
  msq_eval_adj_core(c, x, s, r, gmsq, gf, gx, ga0, ga1, ga2, gm_direct, gw_direct);
  gM += gm_direct + ga0 * c.da_dm[0] + ga1 * c.da_dm[1] + ga2 * c.da_dm[2];
  gW += gw_direct + ga0 * c.da_dw[0] + ga1 * c.da_dw[1] + ga2 * c.da_dw[2];
}

struct CtEval {
  double ct, st, f, tmp, e;
  bool in, lo;
};

GPHS_HD CtEval ct_eval(int kind, double x, double b0, double b1) {
  CtEval r{};
  if (kind == CK_FLAT) {
    r.tmp = 2 * (1 - x);
    r.ct = 1 - r.tmp;
    r.st = math::sqrt_phys(r.tmp * (2 - r.tmp));
    r.f = 1;
    r.in = true;
  } else {
    r.lo = x < 0.5;
    r.e = math::exp_bounded((r.lo ? 2 * x : 2 * (1 - x)) * b1);
    r.tmp = b0 * r.e;
    double c = r.lo ? (r.tmp - b0 - 1) : (-(r.tmp - b0) + 1);
    r.in = c >= -1 && c <= 1;
    double sq = math::sqrt_phys(1 - c * c);
    r.ct = r.in ? c : 1.0;
    r.st = r.in ? sq : 0.0;
    r.f = r.in ? r.tmp * b1 : 0.0;
  }
  return r;
}

// Adjoint of ct_eval: d/dx and d/d(b0, b1).
GPHS_HD void ct_eval_adj(int kind, double x, double b0, double b1, const CtEval& r, double gct, double gst, double gf,
                         double& gx, double& gb0, double& gb1) {
#pragma clang fp contract(fast)
  gx = 0; gb0 = 0; gb1 = 0;
  if (kind == CK_FLAT) {
    // ct = 2x - 1, st = sqrt(tmp (2 - tmp)), tmp = 2 (1 - x)
    gx = 2 * gct;
    if (r.st > 0) gx += gst * math::quot(-2 * (1 - r.tmp), r.st);
    return;
  }
  if (!r.in) return;
  const double sg = r.lo ? 1.0 : -1.0;     // d ct / d tmp
  const double u = r.lo ? x : 1 - x;
  double gc = gct;
  if (r.st > 0) gc += gst * math::quot(-r.ct, r.st);
  const double gtmp = gc * sg + gf * b1;
  gb0 = gc * (-sg) + gtmp * r.e;            // direct -b0 (lo) / +b0 (hi) in ct, tmp = b0 e
  gb1 = gf * r.tmp + gtmp * r.tmp * 2 * u;
  gx = gtmp * r.tmp * 2 * b1 * sg;          // du/dx = sg
}

// Composition M <- M L1 (phs_cpu.cpp:314-340), the same expression order and fused products as
// the forward kernel.
GPHS_HD void compose_l1(double (&M)[16], double g, double bg, double ct, double st, double cp, double sp) {
  const double l01 = -bg * st, l03 = bg * ct;
  const double l11 = ct * cp, l12 = -sp, l13 = st * cp;
  const double l21 = ct * sp, l22 = cp, l23 = st * sp;
  const double l31 = -g * st, l33 = g * ct;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const double r0 = M[4 * i], r1 = M[4 * i + 1], r2 = M[4 * i + 2], r3 = M[4 * i + 3];
    M[4 * i + 0] = __builtin_fma(r3, bg, r0 * g);
    M[4 * i + 1] = __builtin_fma(r3, l31, __builtin_fma(r2, l21, __builtin_fma(r1, l11, r0 * l01)));
    M[4 * i + 2] = __builtin_fma(r2, l22, r1 * l12);
    M[4 * i + 3] = __builtin_fma(r3, l33, __builtin_fma(r2, l23, __builtin_fma(r1, l13, r0 * l03)));
  }
}

// Ma = P L1 out of place, the same operations as compose_l1.
GPHS_HD void compose_l1_to(const double (&P)[16], double (&Ma)[16], double g, double bg, double ct, double st,
                           double cp, double sp) {
  const double l01 = -bg * st, l03 = bg * ct;
  const double l11 = ct * cp, l12 = -sp, l13 = st * cp;
  const double l21 = ct * sp, l22 = cp, l23 = st * sp;
  const double l31 = -g * st, l33 = g * ct;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const double r0 = P[4 * i], r1 = P[4 * i + 1], r2 = P[4 * i + 2], r3 = P[4 * i + 3];
    Ma[4 * i + 0] = __builtin_fma(r3, bg, r0 * g);
    Ma[4 * i + 1] = __builtin_fma(r3, l31, __builtin_fma(r2, l21, __builtin_fma(r1, l11, r0 * l01)));
    Ma[4 * i + 2] = __builtin_fma(r2, l22, r1 * l12);
    Ma[4 * i + 3] = __builtin_fma(r3, l33, __builtin_fma(r2, l23, __builtin_fma(r1, l13, r0 * l03)));
  }
}

// Where the per-event adjoint keeps its tape and stash.  `Store` provides
//   double& T(int node, int k)   L1 of a branch: gamma, bg, ct, st, cp, sp (k = 0..5)
//   double& in(int node, int k)  incoming msq and p_decay of a branch (k = 0, 1)
//   st_put(level, k, v), st_get(level, k): stash level, 16 matrix adjoints (k = 0..15) and 3
//     scalar adjoints (k = 16..18).  A store may drop the matrix entries 2, 4, 8 and 14 (st_get
//     then returns 0) when only the root defers: the stash then only ever meets the root,
//     whose parent boost is the identity, and the adjoint of its L1 reads only the entries at
//     the structural non-zeros of L1.

// Stash slot of entry k in the compact layout (-1: not kept).
GPHS_HD constexpr int compact_stash_slot(int k) {
  return (k == 2 || k == 4 || k == 8 || k == 14) ? -1
         : k < 2 ? k : k < 4 ? k - 1 : k < 8 ? k - 2 : k < 14 ? k - 3 : k - 4;
}


// P = product of the L1 of the ancestors of branch j (root first): the boost its own L1 is
// composed onto.  The ancestors are found by walking the parent links (a few scalar loads, or
// constants in a kernel specialised for the tree's shape).
//   FR: some branches take their polar axis from a beam; their L1 depends on the product of
//   their own ancestors' boosts, which the walk from the root has at hand (identity first).
template <int SID, bool FR, class Store>
GPHS_HD void parent_boost(const NodeCmd* prog, int j, Store& S, double (&P)[16]) {
  using W = Walk<SID>;
  int d = 0;
  for (int a = W::parent(prog, j); a >= 0; a = W::parent(prog, a)) d++;
  if (d == 0 || FR) {
#pragma unroll
    for (int q = 0; q < 16; q++) P[q] = (q % 5 == 0) ? 1.0 : 0.0;
  }
  if (d == 0) return;
  if constexpr (FR) {
    for (int level = 0; level < d; level++) {
      int a = W::parent(prog, j);
      for (int u = 0; u < d - 1 - level; u++) a = W::parent(prog, a);
      if (prog[a].friend_sign != 0)
        compose_friend(P, S.T(a, 0), S.T(a, 1), S.T(a, 2), S.T(a, 3), S.T(a, 4), S.T(a, 5), (double)prog[a].friend_sign);
      else
        compose_l1(P, S.T(a, 0), S.T(a, 1), S.T(a, 2), S.T(a, 3), S.T(a, 4), S.T(a, 5));
    }
    return;
  }
  // the root's L1 itself (the identity times L1: the same values), then the others
  int r = W::parent(prog, j);
  for (int u = 0; u < d - 1; u++) r = W::parent(prog, r);
  {
    const double g = S.T(r, 0), bg = S.T(r, 1), ct = S.T(r, 2), st = S.T(r, 3), cp = S.T(r, 4), sp = S.T(r, 5);
    P[0] = g;   P[1] = -bg * st; P[2] = 0;   P[3] = bg * ct;
    P[4] = 0;   P[5] = ct * cp;  P[6] = -sp; P[7] = st * cp;
    P[8] = 0;   P[9] = ct * sp;  P[10] = cp; P[11] = st * sp;
    P[12] = bg; P[13] = -g * st; P[14] = 0;  P[15] = g * ct;
  }
  for (int level = 1; level < d; level++) {
    int a = W::parent(prog, j);
    for (int u = 0; u < d - 1 - level; u++) a = W::parent(prog, a);
    compose_l1(P, S.T(a, 0), S.T(a, 1), S.T(a, 2), S.T(a, 3), S.T(a, 4), S.T(a, 5));
  }
}

// Adjoint of Ma = P L1 for a branch with a friend (L1 = B_z(bg) (1 (+) Q), Q = R A,
// A = R_z(phi) R_y(theta), R the rotation to the beam's direction n / |n|, n formed from rows 0
// and 3 of P; phs_device.h friend_axis, friend_q).  G: the adjoint of Ma.  Writes the adjoint
// of P into Gc (both terms: G L1^T and the one through n; zero without a parent) and the
// adjoints of the branch's gamma, bg, cos/sin theta and cos/sin phi.
GPHS_HD void friend_adj(const double (&P)[16], const double (&G)[16], bool has_parent, double g, double bg, double ct,
                        double st, double cp, double sp, double f, const FriendAxis& ax, const double (&R)[9],
                        const double (&Q)[9], const double (&L)[16], double (&Gc)[16], double& gg, double& gbg,
                        double& gct, double& gst, double& gcp, double& gsp) {
#pragma clang fp contract(fast)
  double gL[16];
#pragma unroll
  for (int k = 0; k < 4; k++)
#pragma unroll
    for (int j = 0; j < 4; j++)
      gL[4 * k + j] = P[k] * G[j] + P[4 + k] * G[4 + j] + P[8 + k] * G[8 + j] + P[12 + k] * G[12 + j];
  // L1 = [[g, bg q20, bg q21, bg q22], [0, q00 ..], [0, q10 ..], [bg, g q20, g q21, g q22]]
  double gQ[9];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    gQ[j] = gL[5 + j];
    gQ[3 + j] = gL[9 + j];
    gQ[6 + j] = bg * gL[1 + j] + g * gL[13 + j];
  }
  gg = gL[0] + Q[6] * gL[13] + Q[7] * gL[14] + Q[8] * gL[15];
  gbg = gL[12] + Q[6] * gL[1] + Q[7] * gL[2] + Q[8] * gL[3];
  // Q = R A: gA = R^T gQ, gR = gQ A^T; A = [[cp ct, -sp, cp st], [sp ct, cp, sp st], [-st, 0, ct]]
  const double A[9] = {cp * ct, -sp, cp * st, sp * ct, cp, sp * st, -st, 0.0, ct};
  double gA[9], gR[9];
#pragma unroll
  for (int k = 0; k < 3; k++)
#pragma unroll
    for (int j = 0; j < 3; j++) {
      gA[3 * k + j] = R[k] * gQ[j] + R[3 + k] * gQ[3 + j] + R[6 + k] * gQ[6 + j];
      gR[3 * k + j] = gQ[3 * k] * A[3 * j] + gQ[3 * k + 1] * A[3 * j + 1] + gQ[3 * k + 2] * A[3 * j + 2];
    }
  gct = gA[0] * cp + gA[3] * sp + gA[8];
  gst = gA[2] * cp + gA[5] * sp - gA[6];
  gcp = gA[0] * ct + gA[4] + gA[2] * st;
  gsp = -gA[1] + gA[3] * ct + gA[5] * st;
  // hand G L1^T to the parent, then the term through the beam's direction
  if (!has_parent) {
#pragma unroll
    for (int q = 0; q < 16; q++) Gc[q] = 0;
  } else {
#pragma unroll
    for (int ii = 0; ii < 4; ii++)
#pragma unroll
      for (int k = 0; k < 4; k++)
        Gc[4 * ii + k] = G[4 * ii] * L[4 * k] + G[4 * ii + 1] * L[4 * k + 1] + G[4 * ii + 2] * L[4 * k + 2] +
                         G[4 * ii + 3] * L[4 * k + 3];
  }
  // R: r00 = c + h kx^2, r01 = r10 = h kx ky, r11 = c + h ky^2, r02 = -r20 = a0, r12 = -r21 = a1,
  // r22 = c; h = 1 - c, (kx, ky) = (-a1, a0) / sa, sa = |(a0, a1)| (for sa = 0: (1, 0))
  const double a0 = ax.a0, a1 = ax.a1, c = ax.c;
  const double sa = math::sqrt_phys(a0 * a0 + a1 * a1);
  const double kx = sa > 0 ? -a1 / sa : 1.0, ky = sa > 0 ? a0 / sa : 0.0, h = 1 - c;
  const double g01 = gR[1] + gR[3];
  const double gh = gR[0] * kx * kx + g01 * kx * ky + gR[4] * ky * ky;
  const double gkx = 2 * h * kx * gR[0] + h * ky * g01;
  const double gky = 2 * h * ky * gR[4] + h * kx * g01;
  const double gc = gR[0] + gR[4] + gR[8] - gh;
  double ga0 = gR[2] - gR[6], ga1 = gR[5] - gR[7];
  if (sa > 0) {
    ga1 -= gkx / sa;
    ga0 += gky / sa;
    const double gsa = (gkx * a1 - gky * a0) / (sa * sa);
    ga0 += gsa * a0 / sa;
    ga1 += gsa * a1 / sa;
  }
  // (a0, a1, c) = n / |n|
  const double dot = ga0 * a0 + ga1 * a1 + gc * c;
  const double gn0 = (ga0 - a0 * dot) / ax.nn, gn1 = (ga1 - a1 * dot) / ax.nn, gn2 = (gc - c * dot) / ax.nn;
  // n2 = -bg u0 + g u3; u0 = P00 - f P30, n0 = -P01 + f P31, n1 = -P02 + f P32, u3 = -P03 + f P33
  gbg -= ax.u0 * gn2;
  gg += ax.u3 * gn2;
  if (has_parent) {
    const double gu0 = -bg * gn2, gu3 = g * gn2;
    Gc[0] += gu0;  Gc[12] -= f * gu0;
    Gc[1] -= gn0;  Gc[13] += f * gn0;
    Gc[2] -= gn1;  Gc[14] += f * gn1;
    Gc[3] -= gu3;  Gc[15] += f * gu3;
  }
}

// Forward pass of one event with tape recording, then the reverse sweep.
//   xr: the event's n_x randoms (anything indexable); loss_mode 0: cotangents gp [n_out*4] (may be null) and gf
//   are given; loss_mode 1: the loss is factor * sum_j (p_jx^2 + p_jy^2) / s and its
//   cotangents are formed here.  Writes p_out [n_out*4] (may be null), *f_out, *loss_out
//   (mode 1), gx [n_x] (zero for a non-ok event), and adds d/d(M, Gamma) of every branch to
//   gpar[2 * node], gpar[2 * node + 1] (visit index; `GPar` is any indexable accumulator).
//   Returns ok.
//   SID >= 0: the program has that compile-time tree shape (phs_walk.h); both sweeps are then
//   unrolled and every link of the walk is a constant.  LM is the loss mode; with LM = 1 the
//   beams are symmetric (the entry point checks it), so no lab boost, and the cotangents of
//   every leaf's energy and longitudinal momentum are zero, which the sweep leaves out.
//   FR: branches with a friend (nc.friend_sign != 0) take their polar axis from a beam
//   (phs_device.h compose_friend); their L1 depends on their parent boost, whose adjoint then
//   has a second term, through the beam's direction.
// A copy of v that the compiler cannot prove equal to v.  With per-event beams the maps' and
// cos-theta constants depend on the event's sqrt(s); recomputed from such a copy they are
// evaluated again at each use, instead of being kept in registers from the forward pass to
// the end of the reverse sweep (which spilled them to scratch).
GPHS_HD double opaque(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(v));
#endif
  return v;
}
// The same, and not available before `after` has been computed: what is derived from it is
// not scheduled earlier either.
GPHS_HD double opaque_after(double v, double after) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" : "+v"(v) : "v"(after));
#else
  (void)after;
#endif
  return v;
}

template <int SID, int LM, bool PEV, bool FR, class Store, class GPar, class XR>
GPHS_HD bool event_adjoint(const ChannelDesc& cd, const EvBeams& eb, const NodeCmd* prog, const XR& xr,
                           const double* gp, double gf_in, double* p_out, double* f_out, double* loss_out, double* gx,
                           GPar& gpar, Store& S) {
  constexpr int loss_mode = LM;
  using Wk = Walk<SID>;
  const double s = eb.s;
  // a map of daughter c of branch i at this event's beams: the plan's constants, or (PEV) the
  // ones of the event's sqrt(s) with their derivatives with respect to the map's mass and width
  auto child = [&](const NodeCmd& nc, int i, int c) -> ChildCmd {
    if constexpr (!PEV) {
      return nc.k[c];
    } else {
      const double rs = opaque(eb.sqrts);
      ChildCmd r = child_at(nc.k[c], rs, opaque(s), cd.m_tot);
      mc::DualMW a[3], b[3];
      mc::map_constants_d(prog[Wk::child(nc, i, c)].map_id, nc.k[c].m, nc.k[c].w, nc.k[c].m_min, cd.m_tot, rs, a, b);
#pragma unroll
      for (int k = 0; k < 3; k++) { r.da_dm[k] = a[k].dm; r.da_dw[k] = a[k].dw; }
      return r;
    }
  };
  // the event's beams through opaque(): what node_at derives from them is recomputed per use
  auto beams_here = [&]() {
    EvBeams b = eb;
    b.s = opaque(eb.s);
    b.sqrts = opaque(eb.sqrts);
    b.s_friend = opaque(eb.s_friend);
    return b;
  };
  const int N = Wk::n_nodes(cd);
  // ---------------- forward (phs_forward.hip, plus the tape)
  double M[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  double cur_msq = s, cur_m = eb.sqrts, cur_p = 0, F = 1, S2 = 0;
  bool ok = true;
  int depth = 0;
  auto fwd_branch = [&](const int i) {
    const NodeCmd& nc = prog[i];
    const bool leaf[2] = {Wk::leaf(nc, i, 0), Wk::leaf(nc, i, 1)};
    const int push = Wk::push(nc, i), desc = Wk::desc(nc, i);
    S.in(i, 0) = cur_msq; S.in(i, 1) = cur_p;
    double msq[2], f[2], mk[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      if (leaf[c]) { msq[c] = nc.k[c].msq_leaf; f[c] = 1; mk[c] = nc.k[c].m_leaf; }
      else {
        MapEval r;
        if constexpr (PEV) r = msq_eval(child_at(nc.k[c], opaque(eb.sqrts), opaque(s), cd.m_tot), xr[Wk::xi(nc, i, c)], s);
        else r = msq_eval(nc.k[c], xr[Wk::xi(nc, i, c)], s);   // no copy of the record
        msq[c] = r.msq; f[c] = r.f; mk[c] = math::sqrt_phys(r.msq);
      }
    }
    double m_max = nc.m_max, cb0 = nc.b0, cb1 = nc.b1;
    if constexpr (PEV) node_at(nc, i == 0, beams_here(), cd.m_tot, m_max, cb0, cb1);
    const double m = cur_m;
    const double lda = (cur_msq - msq[0] - msq[1]) * (cur_msq - msq[0] - msq[1]) - 4 * msq[0] * msq[1];
    ok = ok && (lda > 0 && m > mk[0] + mk[1] && m <= m_max);
    // quotients to a few ulps (math::quot): the outputs agree with the forward kernel's to
    // ~1e-15 relative, far inside the parity bound
    const double sl = math::sqrt_phys(lda), pk = math::quot(sl, 2 * m);
    F *= f[0] * f[1] * math::quot(sl, cur_msq);
    const double bg = m > 0 ? math::quot(cur_p, m) : 0.0, g = math::sqrt_phys(1 + bg * bg);
    double sp, cp;
    math::sincos_turns(xr[Wk::xphi(nc, i)], sp, cp);
    const CtEval ce = ct_eval(nc.ct_kind, xr[Wk::xphi(nc, i) + 1], cb0, cb1);
    F *= ce.f;
    if (FR && nc.friend_sign != 0) compose_friend(M, g, bg, ce.ct, ce.st, cp, sp, (double)nc.friend_sign);
    else compose_l1(M, g, bg, ce.ct, ce.st, cp, sp);
    S.T(i, 0) = g; S.T(i, 1) = bg; S.T(i, 2) = ce.ct; S.T(i, 3) = ce.st; S.T(i, 4) = cp; S.T(i, 5) = sp;
#pragma unroll
    for (int c = 0; c < 2; c++) {
      if (!leaf[c]) continue;
      const double pc = c == 0 ? pk : -pk;
      const double E = nc.k[c].msq_leaf == 0 ? fabs(pc) : math::sqrt_phys(nc.k[c].msq_leaf + pc * pc);
      double q0 = __builtin_fma(M[3], pc, M[0] * E), q1 = __builtin_fma(M[7], pc, M[4] * E), q2 = __builtin_fma(M[11], pc, M[8] * E),
             q3 = __builtin_fma(M[15], pc, M[12] * E);
      if (LM == 0 && eb.bg_lab != 0) {
        const double b0 = eb.g_lab * q0 + eb.bg_lab * q3, b3 = eb.bg_lab * q0 + eb.g_lab * q3;
        q0 = b0; q3 = b3;
      }
      S2 += q1 * q1 + q2 * q2;
      if (p_out) {
        double* o = p_out + 4 * nc.k[c].out;
        o[0] = q0; o[1] = q1; o[2] = q2; o[3] = q3;
      }
    }
    if (push >= 0) {   // the deferred daughter's kinematics; its boost is re-formed at the pop
      const int j = Wk::child(nc, i, push);
      S.in(j, 0) = msq[push];
      S.in(j, 1) = push == 0 ? pk : -pk;
      depth++;
    }
    if (desc >= 0) {
      cur_msq = msq[desc]; cur_m = mk[desc]; cur_p = desc == 0 ? pk : -pk;
    } else if (depth > 0 && i + 1 < N) {   // the next branch in visit order is the deferred one
      depth--;
      parent_boost<SID, FR>(prog, i + 1, S, M);
      cur_msq = S.in(i + 1, 0); cur_p = S.in(i + 1, 1); cur_m = math::sqrt_phys(cur_msq);
    }
  };
  if constexpr (SID >= 0) {
#pragma unroll
    for (int i = 0; i < Wk::n_nodes(cd); i++) fwd_branch(i);
  } else {
    for (int i = 0; i < N; i++) fwd_branch(i);
  }
  *f_out = F;
  if (loss_out) *loss_out = ok && loss_mode == 1 ? F * S2 / s : 0.0;
  if (!ok) {
    for (int k = 0; k < cd.n_x; k++) gx[k] = 0;
    return false;
  }
  // (an ok event's sweep below writes every entry of gx: one per internal daughter's mass and
  // two per branch's angles are all of the channel's randoms)
  {
  // derivatives only from here on: products and sums may fuse (the outputs above are final)
#pragma clang fp contract(fast)

  // ---------------- cotangents: of the factor here, of the momenta per leaf in the sweep
  const double gF = loss_mode == 1 ? S2 / s : gf_in;
  const double W = gF * F;   // adjoint of log(term) for every factor of the product

  // ---------------- reverse sweep
  double Gc[16] = {0};       // adjoint of the parent boost, from the branch just reversed
  double gc_msq = 0, gc_m = 0, gc_p = 0;
  auto rev_branch = [&](const int i) {
    const NodeCmd& nc = prog[i];
    double m_max_ev, cb0 = nc.b0, cb1 = nc.b1;
    if constexpr (PEV) node_at(nc, i == 0, beams_here(), cd.m_tot, m_max_ev, cb0, cb1);
    const bool leaf[2] = {Wk::leaf(nc, i, 0), Wk::leaf(nc, i, 1)};
    const int push = Wk::push(nc, i), desc = Wk::desc(nc, i);
    double G[16] = {0};
    double g_msq[2] = {0, 0}, g_m[2] = {0, 0}, g_pk = 0;
    if (desc >= 0) {
#pragma unroll
      for (int q = 0; q < 16; q++) G[q] += Gc[q];
      g_msq[desc] += gc_msq; g_m[desc] += gc_m; g_pk += desc == 0 ? gc_p : -gc_p;
    }
    if (push >= 0) {
      depth--;
#pragma unroll
      for (int q = 0; q < 16; q++) G[q] += S.st_get(depth, q);
      g_msq[push] += S.st_get(depth, 16); g_m[push] += S.st_get(depth, 17);
      g_pk += push == 0 ? S.st_get(depth, 18) : -S.st_get(depth, 18);
    }
    // the branch's local quantities: boost parameters from the tape, maps recomputed
    const double a = S.in(i, 0), p0 = S.in(i, 1);
    const double m = i == 0 ? eb.sqrts : math::sqrt_phys(a);
    double P[16];
    parent_boost<SID, FR>(prog, i, S, P);
    MapEval r[2];
    double msq[2], f[2], mk[2], xk[2];
#pragma unroll
    for (int c = 0; c < 2; c++) {
      if (leaf[c]) { msq[c] = nc.k[c].msq_leaf; f[c] = 1; mk[c] = nc.k[c].m_leaf; xk[c] = 0; }
      else {
        xk[c] = xr[Wk::xi(nc, i, c)];
        if constexpr (PEV) r[c] = msq_eval(child_at(nc.k[c], opaque(eb.sqrts), opaque(s), cd.m_tot), xk[c], s);
        else r[c] = msq_eval(nc.k[c], xk[c], s);
        msq[c] = r[c].msq; f[c] = r[c].f; mk[c] = math::sqrt_phys(msq[c]);
      }
    }
    const double d = a - msq[0] - msq[1];
    const double lda = d * d - 4 * msq[0] * msq[1];
    const double sl = math::sqrt_phys(lda), pk = math::quot(sl, 2 * m);
    const double g = S.T(i, 0), bg = S.T(i, 1), cp = S.T(i, 4), sp = S.T(i, 5);
    const double xct = xr[Wk::xphi(nc, i) + 1];
    CtEval ce;
    if (nc.ct_kind == CK_FLAT) {   // from the tape: sin theta as the forward sweep formed it
      ce.tmp = 2 * (1 - xct);
      ce.ct = 1 - ce.tmp;
      ce.st = S.T(i, 3);
      ce.f = 1;
      ce.e = 0;
      ce.in = true;
      ce.lo = false;
    } else {
      ce = ct_eval(nc.ct_kind, xct, cb0, cb1);
    }
    double Ma[16];
    const bool fr = FR && nc.friend_sign != 0;
    // a friend branch: its full L1 = B_z(bg) (1 (+) Q), Q = R(axis) R_z(phi) R_y(theta)
    FriendAxis ax{};
    double fR[9], fQ[9], fL[16];
    if (fr) {
      ax = friend_axis(P, g, bg, (double)nc.friend_sign);
      friend_q(ax, ce.ct, ce.st, cp, sp, fR, fQ);
      fL[0] = g;  fL[1] = bg * fQ[6]; fL[2] = bg * fQ[7]; fL[3] = bg * fQ[8];
      fL[4] = 0;  fL[5] = fQ[0];      fL[6] = fQ[1];      fL[7] = fQ[2];
      fL[8] = 0;  fL[9] = fQ[3];      fL[10] = fQ[4];     fL[11] = fQ[5];
      fL[12] = bg; fL[13] = g * fQ[6]; fL[14] = g * fQ[7]; fL[15] = g * fQ[8];
#pragma unroll
      for (int ii = 0; ii < 4; ii++)
#pragma unroll
        for (int k = 0; k < 4; k++)
          Ma[4 * ii + k] = P[4 * ii] * fL[k] + P[4 * ii + 1] * fL[4 + k] + P[4 * ii + 2] * fL[8 + k] +
                           P[4 * ii + 3] * fL[12 + k];
    } else if (i == 0) {   // the root's parent boost is the identity: Ma = L1
      Ma[0] = g;   Ma[1] = -bg * ce.st; Ma[2] = 0;   Ma[3] = bg * ce.ct;
      Ma[4] = 0;   Ma[5] = ce.ct * cp;  Ma[6] = -sp; Ma[7] = ce.st * cp;
      Ma[8] = 0;   Ma[9] = ce.ct * sp;  Ma[10] = cp; Ma[11] = ce.st * sp;
      Ma[12] = bg; Ma[13] = -g * ce.st; Ma[14] = 0;  Ma[15] = g * ce.ct;
    } else {
      compose_l1_to(P, Ma, g, bg, ce.ct, ce.st, cp, sp);
    }

    // final-state daughters: q = Ma[:,0] E + Ma[:,3] pc
#pragma unroll
    for (int c = 0; c < 2; c++) {
      if (!leaf[c]) continue;
      const double pc = c == 0 ? pk : -pk;
      const double E = nc.k[c].msq_leaf == 0 ? fabs(pc) : math::sqrt_phys(nc.k[c].msq_leaf + pc * pc);
      // cotangent of the lab-frame momentum, moved back through the (linear) lab boost
      double gq[4];
      if (loss_mode == 1) {
        const double q1 = __builtin_fma(Ma[7], pc, Ma[4] * E), q2 = __builtin_fma(Ma[11], pc, Ma[8] * E);
        gq[0] = 0; gq[1] = F * 2 * q1 / s; gq[2] = F * 2 * q2 / s; gq[3] = 0;
      } else if (gp) {
#pragma unroll
        for (int mu = 0; mu < 4; mu++) gq[mu] = gp[4 * nc.k[c].out + mu];
      } else {
        gq[0] = gq[1] = gq[2] = gq[3] = 0;
      }
      if (LM == 0 && eb.bg_lab != 0) {
        const double a0 = gq[0], a3 = gq[3];
        gq[0] = eb.g_lab * a0 + eb.bg_lab * a3;
        gq[3] = eb.bg_lab * a0 + eb.g_lab * a3;
      }
      double gE = 0, gpc = 0;
#pragma unroll
      for (int mu = 0; mu < 4; mu++) {
        if (LM == 1 && (mu == 0 || mu == 3)) continue;   // zero cotangents
        G[4 * mu + 0] = __builtin_fma(gq[mu], E, G[4 * mu + 0]);
        G[4 * mu + 3] = __builtin_fma(gq[mu], pc, G[4 * mu + 3]);
        gE = __builtin_fma(gq[mu], Ma[4 * mu + 0], gE);
        gpc = __builtin_fma(gq[mu], Ma[4 * mu + 3], gpc);
      }
      if (E > 0) gpc += math::quot(gE * pc, E);
      g_pk += c == 0 ? gpc : -gpc;
    }
    double gg, gbg, gct, gst, gcp, gsp;
    if (fr) {
      friend_adj(P, G, i > 0, g, bg, ce.ct, ce.st, cp, sp, (double)nc.friend_sign, ax, fR, fQ, fL, Gc, gg, gbg, gct,
                 gst, gcp, gsp);
    } else {
      // Ma = P L1: gP = G L1^T, gL1 = P^T G (structural non-zeros of L1 only)
      const double l01 = -bg * ce.st, l03 = bg * ce.ct, l11 = ce.ct * cp, l12 = -sp, l13 = ce.st * cp;
      const double l21 = ce.ct * sp, l22 = cp, l23 = ce.st * sp, l31 = -g * ce.st, l33 = g * ce.ct;
      // gP goes straight to Gc, the adjoint handed to the parent (Gc was consumed above)
      if (i > 0) {   // the root has no parent to hand gP to
#pragma unroll
        for (int ii = 0; ii < 4; ii++) {
          const double G0 = G[4 * ii], G1 = G[4 * ii + 1], G2 = G[4 * ii + 2], G3 = G[4 * ii + 3];
          Gc[4 * ii + 0] = __builtin_fma(G3, l03, __builtin_fma(G1, l01, G0 * g));
          Gc[4 * ii + 1] = __builtin_fma(G3, l13, __builtin_fma(G2, l12, G1 * l11));
          Gc[4 * ii + 2] = __builtin_fma(G3, l23, __builtin_fma(G2, l22, G1 * l21));
          Gc[4 * ii + 3] = __builtin_fma(G3, l33, __builtin_fma(G1, l31, G0 * bg));
        }
      } else {
#pragma unroll
        for (int q = 0; q < 16; q++) Gc[q] = 0;
      }
      auto gl = [&](int k, int j) {
        if (i == 0) return G[4 * k + j];   // P = identity
        double acc = P[k] * G[j];
#pragma unroll
        for (int ii = 1; ii < 4; ii++) acc = __builtin_fma(P[4 * ii + k], G[4 * ii + j], acc);
        return acc;
      };
      const double g00 = gl(0, 0), g01 = gl(0, 1), g03 = gl(0, 3), g11 = gl(1, 1), g12 = gl(1, 2), g13 = gl(1, 3);
      const double g21 = gl(2, 1), g22 = gl(2, 2), g23 = gl(2, 3), g30 = gl(3, 0), g31 = gl(3, 1), g33 = gl(3, 3);
      gg = g00 - g31 * ce.st + g33 * ce.ct;
      gbg = -g01 * ce.st + g03 * ce.ct + g30;
      gct = g03 * bg + g11 * cp + g21 * sp + g33 * g;
      gst = -g01 * bg + g13 * cp + g23 * sp - g31 * g;
      gcp = g11 * ce.ct + g13 * ce.st + g22;
      gsp = -g12 + g21 * ce.ct + g23 * ce.st;
    }
    // phi = 2 pi x
    gx[Wk::xphi(nc, i)] = (gsp * cp - gcp * sp) * kTwoPi;
    // cos theta map of this branch (its own M through b0, b1)
    const double gfct = ce.f != 0 ? math::quot(W, ce.f) : 0.0;
    double gxc, gb0, gb1;
    ct_eval_adj(nc.ct_kind, xct, cb0, cb1, ce, gct, gst, gfct, gxc, gb0, gb1);
    gx[Wk::xphi(nc, i) + 1] = gxc;
    if constexpr (PEV) {   // the cos-theta constants at this event's s: b0 = M^2 / s, b1 = log((b0 + 1) / b0)
      double db0 = 0, db1 = 0;
      if (nc.ct_kind == CK_PEAKED) {
        mc::DualMW b[2];
        mc::ct_constants_d(nc.m, nc.friend_sign != 0 ? opaque(eb.s_friend) : opaque(eb.s), b);
        db0 = b[0].dm;
        db1 = b[1].dm;
      }
      gpar[2 * i] += gb0 * db0 + gb1 * db1;
    } else {
      gpar[2 * i] += gb0 * nc.db_dm[0] + gb1 * nc.db_dm[1];
    }
    // gamma = sqrt(1 + bg^2), bg = p / m
    gbg += math::quot(gg * bg, g);
    double g_a = 0, g_mnode = 0, g_p0 = 0;
    const double inv_m = m > 0 ? math::quot(1.0, m) : 0.0;
    if (m > 0) { g_p0 += gbg * inv_m; g_mnode -= gbg * p0 * inv_m * inv_m; }
    // pk = sl / (2 m)
    double gsl = g_pk * inv_m * 0.5;
    g_mnode -= g_pk * sl * inv_m * inv_m * 0.5;
    // factor term f0 f1 sl / a
    double gfk[2] = {0, 0};
    if (W != 0) {
#pragma unroll
      for (int c = 0; c < 2; c++)
        if (f[c] != 0) gfk[c] = math::quot(W, f[c]);
      if (sl != 0) gsl += math::quot(W, sl);
      g_a -= math::quot(W, a);
    }
    // lambda = d^2 - 4 msq0 msq1, d = a - msq0 - msq1
    const double glda = sl > 0 ? math::quot(gsl, 2 * sl) : 0.0;
    const double gd = 2 * d * glda;
    g_msq[0] += -4 * msq[1] * glda - gd;
    g_msq[1] += -4 * msq[0] * glda - gd;
    g_a += gd;
    // internal daughters: m_k = sqrt(msq_k), then their msq maps
#pragma unroll
    for (int c = 0; c < 2; c++) {
      if (leaf[c]) continue;
      if (mk[c] > 0) g_msq[c] += math::quot(g_m[c], 2 * mk[c]);
      double gxk = 0, gM = 0, gW = 0;
      if constexpr (PEV && SID >= 0) {
        // the derivatives of the map constants at this event's sqrt(s) wait for the end of
        // the sweep, where fewer registers are live; the adjoints of the constants wait in
        // this branch's tape entries, which nothing reads after this point
        double ga[3];
        msq_eval_adj_core(child_at(nc.k[c], opaque(eb.sqrts), opaque(s), cd.m_tot), xk[c], s, r[c], g_msq[c],
                          gfk[c], gxk, ga[0], ga[1], ga[2], gM, gW);
        S.T(i, 3 * c) = ga[0]; S.T(i, 3 * c + 1) = ga[1]; S.T(i, 3 * c + 2) = ga[2];
      } else if constexpr (PEV) {
        msq_eval_adj(child(nc, i, c), xk[c], s, r[c], g_msq[c], gfk[c], gxk, gM, gW);
      } else {
        msq_eval_adj(nc.k[c], xk[c], s, r[c], g_msq[c], gfk[c], gxk, gM, gW);
      }
      gx[Wk::xi(nc, i, c)] = gxk;
      gpar[2 * Wk::child(nc, i, c)] += gM;
      gpar[2 * Wk::child(nc, i, c) + 1] += gW;
    }
    // hand the adjoints of this branch's inputs to its parent
    gc_msq = g_a; gc_m = g_mnode; gc_p = g_p0;
    if (i >= 1 && Wk::desc(prog[i - 1], i - 1) < 0) {   // reached by a pop in the forward pass
#pragma unroll
      for (int q = 0; q < 16; q++) S.st_put(depth, q, Gc[q]);
      S.st_put(depth, 16, gc_msq); S.st_put(depth, 17, gc_m); S.st_put(depth, 18, gc_p);
      depth++;
    }
  };
  if constexpr (SID >= 0) {
#pragma unroll
    for (int i = Wk::n_nodes(cd) - 1; i >= 0; i--) rev_branch(i);
    if constexpr (PEV) {
      // d a[k] / d (M, Gamma) of every internal daughter's map at this event's sqrt(s), one
      // daughter after the other (each waits for the previous one's result)
      double chain = 0;
#if defined(__HIP_DEVICE_COMPILE__)
      asm volatile("" ::: "memory");   // the adjoints are read back from the tape, not kept in registers
#endif
#pragma unroll
      for (int i = 0; i < Wk::n_nodes(cd); i++) {
        const NodeCmd& nc = prog[i];
#pragma unroll
        for (int c = 0; c < 2; c++) {
          if (Wk::leaf(nc, i, c)) continue;
          const int j = Wk::child(nc, i, c);
          const double ga0 = S.T(i, 3 * c), ga1 = S.T(i, 3 * c + 1), ga2 = S.T(i, 3 * c + 2);
          // a[0] of the Breit-Wigner map is the plan's (the map itself uses the plan's a[0])
          const mc::DualMW a0_bw(nc.k[c].a0, nc.k[c].da_dm[0], nc.k[c].da_dw[0]);
          mc::DualMW a[3], b[3];
          mc::map_constants_d(prog[j].map_id, nc.k[c].m, nc.k[c].w, nc.k[c].m_min, cd.m_tot,
                              opaque_after(opaque_after(eb.sqrts, ga0), chain), a, b, 0.0, &a0_bw);
          gpar[2 * j] += ga0 * a[0].dm + ga1 * a[1].dm + ga2 * a[2].dm;
          gpar[2 * j + 1] += ga0 * a[0].dw + ga1 * a[1].dw + ga2 * a[2].dw;
          chain = gpar[2 * j + 1];
        }
      }
    }
  } else {
    for (int i = N - 1; i >= 0; i--) rev_branch(i);
  }
  }
  return true;
}

}  // namespace gphs
