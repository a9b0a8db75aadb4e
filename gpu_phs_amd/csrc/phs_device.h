// Device-side building blocks shared by the forward and adjoint kernels.
//
// The arithmetic follows the reference's serial algorithm expression by expression
// (phs_cpu.cpp:12-61 msq maps, 112-146 cos-theta maps, 174-234 decay, 283-345 angles and
// boosts) so that, with contraction off, results differ from the CPU only where a
// transcendental (sin, cos, tan, exp, log: phs_math.h) rounds differently.  Where the
// reference computes a value and then branches on x (T/U-channel maps, peaked cos-theta map),
// the kernel evaluates one transcendental on a selected argument and selects the result, so
// lanes never diverge.
#pragma once
#include <hip/hip_runtime.h>

#include "phs_math.h"
#include "phs_types.h"

namespace gphs {

// Pointers read from the segment table are generic; casting them to the global address space
// lets the compiler emit global_* instead of flat_* memory instructions.
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) uint8_t guint8;
// 16-byte words of global memory (two doubles)
typedef double dvec2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) dvec2 gdouble2;
__device__ __forceinline__ gdouble* as_global(double* p) { return (gdouble*)p; }
__device__ __forceinline__ const gdouble* as_global(const double* p) { return (const gdouble*)p; }
__device__ __forceinline__ guint8* as_global(uint8_t* p) { return (guint8*)p; }

#define GPHS_PI 3.14159265358979323846

// msq of an internal daughter and its mapping factor (phs_cpu.cpp:12-61).
__device__ __forceinline__ void msq_map(const ChildCmd& c, double x, double s, double& msq, double& f) {
  switch (c.kind) {
    case MK_LINEAR:
      msq = (1 - x) * c.msq_min + x * c.msq_max;
      f = c.f_lin;
      break;
    case MK_BREIT_WIGNER: {
      double z = (1 - x) * c.a0 + x * c.a1;
      double t = math::tan_radians(z);
      bool in = (-GPHS_PI / 2 < z) && (z < GPHS_PI / 2);
      msq = in ? c.m * (c.m + c.w * t) : 0.0;
      f = in ? c.a2 * (1 + t * t) : 0.0;
      break;
    }
    case MK_EXP: {
      double m1 = c.a0 * math::exp_bounded(x * c.a1);
      msq = m1 - c.a0 + c.msq_min;
      f = c.a2 * m1;
      break;
    }
    case MK_TU: {
      bool lo = x < 0.5;
      double m1 = c.a0 * math::exp_bounded((lo ? x : (1 - x)) * c.a1);
      msq = lo ? (m1 - c.a0 + c.msq_min) : (-(m1 - c.a0) + c.msq_max);
      f = c.a2 * m1;
      break;
    }
    case MK_STEP_H: {
      double z = c.a1 / (c.a0 - x) - c.a1 / c.a0 + c.a2 * x;
      msq = z * c.msq_max + (1 - z) * c.msq_min;
      f = (c.a1 / ((c.a0 - x) * (c.a0 - x)) + c.a2) * (c.msq_max - c.msq_min) / s;
      break;
    }
    default: {  // MK_STEP_E
      double tmp = math::exp_bounded(-x * c.a0 / c.a2) * (1 + c.a1);
      double z = -c.a2 * math::log_bounded(tmp - c.a1);
      msq = z * c.msq_max + (1 - z) * c.msq_min;
      f = c.a0 / (1 - c.a1 / tmp) * (c.msq_max - c.msq_min) / s;
      break;
    }
  }
}

// The beams of one event: the channel's fixed ones, or the event's own energies E1, E2
// (phs.cu:1313-1317: sqrt(s) = 2 sqrt(E1 E2); phs.cu:1289: bg_lab = (E1 - E2) / (2 sqrt(E1 E2));
// phs.cu:879: a branch with a friend evaluates its cos-theta map at s = 4 E1 E2).
struct EvBeams {
  double sqrts, s, bg_lab, g_lab, s_friend, volume;
};

__device__ __forceinline__ EvBeams fixed_beams(const ChannelDesc& cd) {
  return EvBeams{cd.sqrts, cd.s, cd.bg_lab, cd.g_lab, cd.s, cd.volume};
}

#define GPHS_TWOPI2 39.47841760435743447534   // mappings.h:6
#define GPHS_TWOPI5 9792.62991312900650440772 // mappings.h:7

__device__ __forceinline__ EvBeams event_beams(const ChannelDesc& cd, double E1, double E2) {
  EvBeams b;
  const double r = math::sqrt_phys(E1 * E2);
  b.sqrts = 2 * r;
  b.s = b.sqrts * b.sqrts;
  b.bg_lab = (E1 - E2) / (2 * r);
  b.g_lab = math::sqrt_phys(1 + b.bg_lab * b.bg_lab);
  b.s_friend = 4 * E1 * E2;
  // the phase-space volume (phs_cpu.cpp:195-212): s / (4 (2 pi)^2) per internal branch below
  // the root, 1 / (4 (2 pi)^5) for the root
  double v = 1;
  for (int i = 1; i < cd.n_nodes; i++) v = v * b.sqrts * b.sqrts / (4 * GPHS_TWOPI2);
  b.volume = v / (4 * GPHS_TWOPI5);
  return b;
}

// An internal daughter's map at this event's sqrt(s): the constants that depend on it,
// recomputed with the expressions of init_mapping_constants_cpu (phs_cpu.cpp:452-511);
// a0 of the Breit-Wigner, exponential and t/u maps does not depend on sqrt(s).
__device__ __forceinline__ ChildCmd child_at(const ChildCmd& c, double sqrts, double s, double m_tot) {
  ChildCmd r = c;
  if (c.leaf) return r;
  const double m_max = sqrts - m_tot + c.m_min;
  const double msq_min = c.msq_min, msq_max = m_max * m_max, d = msq_max - msq_min;
  r.msq_max = msq_max;
  r.f_lin = d / s;
  switch (c.kind) {
    case MK_LINEAR: r.a1 = d; r.a2 = r.a1 / s; break;
    case MK_BREIT_WIGNER: {
      const double msq0 = c.m * c.m;
      r.a1 = math::atan_radians((msq_max - msq0) / (c.m * c.w));
      r.a2 = (r.a1 - c.a0) * (c.m * c.w) / s;
      break;
    }
    case MK_EXP: r.a1 = math::log_bounded(d / c.a0 + 1); r.a2 = r.a1 / s; break;
    case MK_TU: r.a1 = 2 * math::log_bounded(d / (2 * c.a0) + 1); r.a2 = r.a1 / s; break;
    case MK_STEP_E: {
      double a2 = 2 * c.m * c.w / d;
      a2 = a2 < 0.01 ? 0.01 : a2;
      const double a1 = math::exp_bounded(-(c.m * c.m - msq_min) / d / a2);
      r.a2 = a2;
      r.a1 = a1;
      r.a0 = 1 - a2 * math::log_bounded((1 + a1 * math::exp_bounded(1 / a2)) / (1 + a1));
      break;
    }
    default: {  // MK_STEP_H
      const double a2 = (c.m * c.m - msq_min) / d;
      const double q = 2 * c.m * c.w / d;
      double a1 = q * q / a2;
      a1 = a1 < 0.000001 ? 0.000001 : a1;
      r.a2 = a2;
      r.a1 = a1;
      r.a0 = (1 + math::sqrt_phys(1 + 4 * a1 / (1 - a2))) / 2;
      break;
    }
  }
  return r;
}

// A branch's kinematic bound and cos-theta constants at this event's beams
// (phs_cpu.cpp:514-527; mappings_gpu.hpp:186-188 at s = 4 E1 E2 for a friend).
__device__ __forceinline__ void node_at(const NodeCmd& nc, bool root, const EvBeams& b, double m_tot, double& m_max,
                                        double& b0, double& b1) {
  m_max = root ? b.sqrts : b.sqrts - m_tot + nc.m_min;
  if (nc.ct_kind == CK_PEAKED) {
    b0 = nc.m * nc.m / (nc.friend_sign != 0 ? b.s_friend : b.s);
    b1 = math::log_bounded((b0 + 1) / b0);
  } else {
    b0 = nc.b0;
    b1 = nc.b1;
  }
}

// cos theta, sin theta and the mapping factor (phs_cpu.cpp:112-146).
__device__ __forceinline__ void ct_map(int kind, double x, double b0, double b1, double& ct, double& st, double& f) {
  if (kind == CK_FLAT) {
    double tmp = 2 * (1 - x);
    ct = 1 - tmp;
    st = math::sqrt_phys(tmp * (2 - tmp));
    f = 1;
  } else {
    bool lo = x < 0.5;
    double tmp = b0 * math::exp_bounded((lo ? 2 * x : 2 * (1 - x)) * b1);
    double c = lo ? (tmp - b0 - 1) : (-(tmp - b0) + 1);
    bool in = c >= -1 && c <= 1;
    double sq = math::sqrt_phys(1 - c * c);
    ct = in ? c : 1.0;
    st = in ? sq : 0.0;
    f = in ? tmp * b1 : 0.0;
  }
}

// M <- M * L1 with L1 = B_z(bg) R_z(phi) R_y(theta) (phs_cpu.cpp:314-340); the products with
// the structural zeros of L1 are left out, the others are summed in the reference's order,
// each after the first as a fused multiply-add (one rounding per term instead of two).
__device__ __forceinline__ void compose(double (&M)[16], double g, double bg, double ct, double st, double cp, double sp) {
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

// Columns 0 and 3 of M * L1 only: all that the momenta of two final-state daughters need
// (p = L (E, 0, 0, +-p)); columns 1 and 2 are left stale, so M must not be used further.
__device__ __forceinline__ void compose_cols03(double (&M)[16], double g, double bg, double ct, double st, double cp,
                                               double sp) {
  const double l03 = bg * ct, l13 = st * cp, l23 = st * sp, l33 = g * ct;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const double r0 = M[4 * i], r1 = M[4 * i + 1], r2 = M[4 * i + 2], r3 = M[4 * i + 3];
    M[4 * i + 0] = __builtin_fma(r3, bg, r0 * g);
    M[4 * i + 3] = __builtin_fma(r3, l33, __builtin_fma(r2, l23, __builtin_fma(r1, l13, r0 * l03)));
  }
}

// The beam's direction in the branch's rest frame, n / |n| = (a0, a1, c), with the
// intermediates the adjoint needs (u0 and u3: the beam's energy and z component seen from the
// parent frame, n2 = -bg u0 + g u3).
struct FriendAxis {
  double u0, u3, n0, n1, n2, nn, a0, a1, c;
};
__host__ __device__ __forceinline__ FriendAxis friend_axis(const double (&M)[16], double g, double bg, double f) {
  FriendAxis x;
  x.u0 = M[0] - f * M[12];
  x.n0 = -M[1] + f * M[13];
  x.n1 = -M[2] + f * M[14];
  x.u3 = -M[3] + f * M[15];
  x.n2 = -bg * x.u0 + g * x.u3;
  x.nn = math::sqrt_phys(x.n0 * x.n0 + x.n1 * x.n1 + x.n2 * x.n2);
  x.a0 = x.n0 / x.nn;
  x.a1 = x.n1 / x.nn;
  x.c = x.n2 / x.nn;
  return x;
}

// R (Rodrigues, about e_z x n, row-major r[3 a + b]) and Q = R R_z(phi) R_y(theta).
__host__ __device__ __forceinline__ void friend_q(const FriendAxis& ax, double ct, double st, double cp, double sp,
                                                  double (&r)[9], double (&q)[9]) {
  const double a0 = ax.a0, a1 = ax.a1, c = ax.c;
  const double sa = math::sqrt_phys(a0 * a0 + a1 * a1);
  const double kx = sa > 0 ? -a1 / sa : 1.0;
  const double ky = sa > 0 ? a0 / sa : 0.0;
  const double h = 1 - c;
  r[0] = c + h * kx * kx; r[1] = h * kx * ky; r[2] = a0;
  r[3] = r[1];            r[4] = c + h * ky * ky; r[5] = a1;
  r[6] = -a0;             r[7] = -a1;             r[8] = c;
  const double rc0 = cp * ct, rc1 = sp * ct, rs0 = cp * st, rs1 = sp * st;
#pragma unroll
  for (int a = 0; a < 3; a++) {
    q[3 * a + 0] = r[3 * a] * rc0 + r[3 * a + 1] * rc1 + r[3 * a + 2] * -st;
    q[3 * a + 1] = r[3 * a] * -sp + r[3 * a + 1] * cp;
    q[3 * a + 2] = r[3 * a] * rs0 + r[3 * a + 1] * rs1 + r[3 * a + 2] * ct;
  }
}

// M <- M * L1 for a branch whose polar axis is an incoming beam, its "friend"
// (phs.cu:853-910).  The beam k = (1, 0, 0, f) (f = +1 along +z, -1 along -z; the reference
// scales k by the beam energy, which the normalisation removes) is seen from the branch's
// rest frame as n = spatial part of B_z(bg)^-1 M^-1 k, with M^-1 = eta M^T eta (friend_axis).
// R is the rotation about e_z x n that takes e_z to n / |n| (Rodrigues' formula; for n along
// -e_z, the half turn about e_x), and L1 = B_z(bg) (1 (+) R R_z(phi) R_y(theta)), so the
// first daughter's polar angle theta is measured from the beam.  The reference's matrix for
// R (L_rot_to_2nd, phs.cu:753-781) is not orthogonal and the bg term of its n_z has the
// opposite sign; this is the rotation those lines describe.  Every product and sum is in
// the order of the CPU restatement, which forms the same matrices.
__device__ __forceinline__ void compose_friend(double (&M)[16], double g, double bg, double ct, double st, double cp,
                                               double sp, double f) {
  double r[9], q[9];
  friend_q(friend_axis(M, g, bg, f), ct, st, cp, sp, r, q);
  const double l01 = bg * q[6], l02 = bg * q[7], l03 = bg * q[8];
  const double l31 = g * q[6], l32 = g * q[7], l33 = g * q[8];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const double x0 = M[4 * i], x1 = M[4 * i + 1], x2 = M[4 * i + 2], x3 = M[4 * i + 3];
    M[4 * i + 0] = x0 * g + x3 * bg;
    M[4 * i + 1] = x0 * l01 + x1 * q[0] + x2 * q[3] + x3 * l31;
    M[4 * i + 2] = x0 * l02 + x1 * q[1] + x2 * q[4] + x3 * l32;
    M[4 * i + 3] = x0 * l03 + x1 * q[2] + x2 * q[5] + x3 * l33;
  }
}

// L1 itself (the root's parent boost is the identity).
__device__ __forceinline__ void set_l1(double (&M)[16], double g, double bg, double ct, double st, double cp, double sp) {
  M[0] = g;   M[1] = -bg * st; M[2] = 0;   M[3] = bg * ct;
  M[4] = 0;   M[5] = ct * cp;  M[6] = -sp; M[7] = st * cp;
  M[8] = 0;   M[9] = ct * sp;  M[10] = cp; M[11] = st * sp;
  M[12] = bg; M[13] = -g * st; M[14] = 0;  M[15] = g * ct;
}

}  // namespace gphs
