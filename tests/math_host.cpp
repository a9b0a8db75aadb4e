// Host build of the kernels' bounded-range elementary functions (gpu_phs_amd/csrc/phs_math.h), so the
// CPU tests can check it against high-precision arithmetic.
#include "phs_math.h"

extern "C" {
void host_sincos_turns(const double* x, long n, double* s, double* c) {
  for (long i = 0; i < n; i++) gphs::math::sincos_turns(x[i], s[i], c[i]);
}
void host_tan_radians(const double* z, long n, double* t) {
  for (long i = 0; i < n; i++) t[i] = gphs::math::tan_radians(z[i]);
}
void host_exp(const double* x, long n, double* y) {
  for (long i = 0; i < n; i++) y[i] = gphs::math::exp_bounded(x[i]);
}
void host_log(const double* x, long n, double* y) {
  for (long i = 0; i < n; i++) y[i] = gphs::math::log_bounded(x[i]);
}
void host_atan(const double* x, long n, double* y) {
  for (long i = 0; i < n; i++) y[i] = gphs::math::atan_radians(x[i]);
}
void host_atan2_turns(const double* y, const double* x, long n, double* t) {
  for (long i = 0; i < n; i++) t[i] = gphs::math::atan2_turns(y[i], x[i]);
}
}
