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
// the structural zeros of L1 are left out, every other product and sum keeps the reference's
// order.
__device__ __forceinline__ void compose(double (&M)[16], double g, double bg, double ct, double st, double cp, double sp) {
  const double l01 = -bg * st, l03 = bg * ct;
  const double l11 = ct * cp, l12 = -sp, l13 = st * cp;
  const double l21 = ct * sp, l22 = cp, l23 = st * sp;
  const double l31 = -g * st, l33 = g * ct;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const double r0 = M[4 * i], r1 = M[4 * i + 1], r2 = M[4 * i + 2], r3 = M[4 * i + 3];
    M[4 * i + 0] = r0 * g + r3 * bg;
    M[4 * i + 1] = r0 * l01 + r1 * l11 + r2 * l21 + r3 * l31;
    M[4 * i + 2] = r1 * l12 + r2 * l22;
    M[4 * i + 3] = r0 * l03 + r1 * l13 + r2 * l23 + r3 * l33;
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
    M[4 * i + 0] = r0 * g + r3 * bg;
    M[4 * i + 3] = r0 * l03 + r1 * l13 + r2 * l23 + r3 * l33;
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
