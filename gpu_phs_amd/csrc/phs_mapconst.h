// The map constants of one branch and their derivatives with respect to the map's mass and
// width, shared by the host compiler (tree_compiler.cpp: every channel's constants and Jacobians
// at the process's sqrt(s)) and the adjoint kernel (the Jacobians at each event's own sqrt(s)
// when the beams vary per event).
#pragma once
#include <cmath>

#include "phs_math.h"
#include "phs_types.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GPHS_MC __host__ __device__ inline
#define GPHS_MC_CTOR __host__ __device__
#else
#define GPHS_MC inline
#define GPHS_MC_CTOR
#endif

namespace gphs {
// In a namespace of their own: the overloads of atan, log, exp and sqrt below take DualMW, and
// a double converts to DualMW, so they must not be found by unqualified calls elsewhere.
namespace mc {

// Value and derivatives with respect to (M, Gamma) of a map constant: lets the same code
// give the constants and their Jacobians, which the adjoint kernel chains through.
struct DualMW {
  double v = 0, dm = 0, dw = 0;
  GPHS_MC_CTOR DualMW() = default;
  GPHS_MC_CTOR DualMW(double x) : v(x) {}
  GPHS_MC_CTOR DualMW(double x, double a, double b) : v(x), dm(a), dw(b) {}
};
GPHS_MC DualMW operator+(DualMW a, DualMW b) { return {a.v + b.v, a.dm + b.dm, a.dw + b.dw}; }
GPHS_MC DualMW operator-(DualMW a, DualMW b) { return {a.v - b.v, a.dm - b.dm, a.dw - b.dw}; }
GPHS_MC DualMW operator-(DualMW a) { return {-a.v, -a.dm, -a.dw}; }
GPHS_MC DualMW operator*(DualMW a, DualMW b) { return {a.v * b.v, a.dm * b.v + a.v * b.dm, a.dw * b.v + a.v * b.dw}; }
GPHS_MC DualMW operator/(DualMW a, DualMW b) {
  return {a.v / b.v, (a.dm * b.v - a.v * b.dm) / (b.v * b.v), (a.dw * b.v - a.v * b.dw) / (b.v * b.v)};
}
GPHS_MC DualMW d2chain(DualMW a, double f, double df) { return {f, df * a.dm, df * a.dw}; }
// On the host, the C library's functions, as the reference calls them; in the kernels (the
// constants at each event's sqrt(s)), the library's own of phs_math.h, whose constants are formed
// where they are used (the device library's keep theirs in vector registers, which the adjoint
// kernel then spills).  pow(a, 2) is a * a, correctly rounded either way.
#if defined(__HIP_DEVICE_COMPILE__)
GPHS_MC double f_atan(double x) { return math::atan_radians(x); }
GPHS_MC double f_log(double x) { return math::log_bounded(x); }
GPHS_MC double f_exp(double x) { return math::exp_bounded(x); }
GPHS_MC double f_sqrt(double x) { return math::sqrt_phys(x); }
GPHS_MC double f_sq(double x) { return x * x; }
#else
GPHS_MC double f_atan(double x) { return std::atan(x); }
GPHS_MC double f_log(double x) { return std::log(x); }
GPHS_MC double f_exp(double x) { return std::exp(x); }
GPHS_MC double f_sqrt(double x) { return std::sqrt(x); }
GPHS_MC double f_sq(double x) { return std::pow(x, 2); }
#endif
GPHS_MC DualMW atan(DualMW a) { return d2chain(a, f_atan(a.v), 1 / (1 + a.v * a.v)); }
GPHS_MC DualMW log(DualMW a) { return d2chain(a, f_log(a.v), 1 / a.v); }
GPHS_MC DualMW exp(DualMW a) { double e = f_exp(a.v); return d2chain(a, e, e); }
GPHS_MC DualMW sqrt(DualMW a) { double r = f_sqrt(a.v); return d2chain(a, r, 0.5 / r); }
GPHS_MC DualMW pow2(DualMW a) { return d2chain(a, f_sq(a.v), 2 * a.v); }   // pow(a, 2) of the reference
GPHS_MC DualMW dmax(DualMW a, double b) { return a.v < b ? DualMW(b) : a; }           // std::max(a, b)
GPHS_MC double val(const DualMW& a) { return a.v; }

// The cos-theta constants of a branch whose cos-theta map is peaked (the collinear, t- and
// u-channel maps) at s = sb: b0 = M^2 / sb, b1 = log((b0 + 1) / b0), phs_cpu.cpp:515-520.
GPHS_MC void ct_constants_d(double m_, double sb, DualMW b[2]) {
  const DualMW m(m_, 1, 0);
  b[0] = m * m / sb;
  b[1] = log((b[0] + 1) / b[0]);
}

// phs_cpu.cpp:452-531 (init_mapping_constants_cpu) for one branch, same expressions and
// evaluation order; the values are those of the reference, bit for bit, and dm / dw carry
// their derivatives with respect to the map mass and width.  s_ct, when nonzero, is the s the
// cos-theta constants are taken at (a branch with a friend: 4 E1 E2, phs.cu:879).
//
// a0_bw, when given, is the Breit-Wigner map's a[0] with its derivatives, which do not depend
// on sqrt(s): the constant is then taken as it is instead of being formed again.
GPHS_MC void map_constants_d(int id, double m_, double w_, double m_min, double m_tot, double sqrts, DualMW a[3], DualMW b[3],
                     double s_ct = 0, const DualMW* a0_bw = nullptr) {
  const DualMW m(m_, 1, 0), w(w_, 0, 1);
  double m_max = sqrts - m_tot + m_min;
  double msq_min = m_min * m_min, msq_max = m_max * m_max, s = sqrts * sqrts;
  const double sb = s_ct != 0 ? s_ct : s;
  DualMW msq0;
  a[0] = a[1] = a[2] = DualMW(0);
  switch (id) {
    case MAP_NO:
      a[0] = 0; a[1] = msq_max - msq_min; a[2] = a[1] / s; break;
    case MAP_SCHANNEL:
      msq0 = m * m;
      a[0] = a0_bw ? *a0_bw : atan((msq_min - msq0) / (m * w));
      a[1] = atan((msq_max - msq0) / (m * w));
      a[2] = (a[1] - a[0]) * (m * w) / s; break;
    case MAP_RADIATION: case MAP_COLLINEAR: case MAP_INFRARED:
      msq0 = id == MAP_RADIATION ? w * w : m * m;
      a[0] = msq0;
      a[1] = log((msq_max - msq_min) / msq0 + 1);
      a[2] = a[1] / s; break;
    case MAP_TCHANNEL: case MAP_UCHANNEL:
      msq0 = m * m;
      a[0] = msq0;
      a[1] = 2 * log((msq_max - msq_min) / (2 * msq0) + 1);
      a[2] = a[1] / s; break;
    case MAP_STEP_E:
      a[2] = dmax(2 * m * w / (msq_max - msq_min), 0.01);
      a[1] = exp(-(m * m - msq_min) / (msq_max - msq_min) / a[2]);
      a[0] = 1 - a[2] * log((1 + a[1] * exp(1 / a[2])) / (1 + a[1])); break;
    case MAP_STEP_H:
      a[2] = (m * m - msq_min) / (msq_max - msq_min);
      a[1] = dmax(pow2(2 * m * w / (msq_max - msq_min)) / a[2], 0.000001);
      a[0] = (1 + sqrt(1 + 4 * a[1] / (1 - a[2]))) / 2; break;
    default: break;
  }
  switch (id) {
    case MAP_TCHANNEL: case MAP_UCHANNEL: case MAP_COLLINEAR:
      ct_constants_d(m_, sb, b); b[2] = 0; break;
    default:
      b[0] = b[1] = b[2] = DualMW(0); break;
  }
}

}  // namespace mc
}  // namespace gphs
