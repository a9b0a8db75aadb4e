// ORACLE — test infrastructure only.
//
// CPU restatement of the reference's serial phase-space algorithm (phs_cpu.cpp,
// Whizard's recursion).  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker: the product
// path (gpu_phs_amd, libgpu_phs.so) never calls it.
//
// Every function cites the reference file:line it restates.  The scalar type is a
// template parameter so that the same code runs in double (the oracle proper) and
// in forward-mode dual numbers (the derivative reference for the adjoint kernel).
//
// Pinning (what checks this restatement, see tests/test_oracle.py):
//   * bitwise against the compiled reference phs_cpu.cpp (oracle/_ref) on seeded
//     randoms for every map id and for the ttH / bwbw trees;
//   * Whizard's own momenta in samples/ttH.ref and samples/bwbw.ref (1e-6, the
//     randoms there carry 9 decimals);
//   * closed forms: 2->2 |p| = sqrt(lambda(s,m^2,m^2))/(2 sqrt s), volume
//     s^(n-2)/(4^(n-1) (2 pi)^(2n+1)), map end points, f_map = (d msq/dx)/s and
//     f_ct = (d ct/dx)/2 by finite differences, momentum conservation, on-shell
//     masses;
//   * the dual-number path against central finite differences of the double path.
//   Parts with no counterpart in phs_cpu.cpp:
//   * the per-event sqrt(s) (reduces to the fixed-sqrt(s) path when E1 = E2) and the
//     lab-frame boost (phs.cu:1286-1301; identity for E1 = E2, Lorentz invariants and
//     total momentum otherwise), pinned by those invariants;
//   * the friend axis (phs.cu:853-910; GPU-only in the reference, whose rotation matrix
//     there is not orthogonal): Whizard's momenta and all-channel factors of the friend
//     trees of samples/bwbw.ref, the polar angle of the first daughter measured from the
//     beam in the branch's rest frame, conservation, on-shell masses, and the f = +1 root
//     reducing to the plain composition;
//   * the inverse mapping (phs.cu:626-675, 924-1284): the round trip into the generating
//     channel and Whizard's all-channel factors of samples/ttH.ref and samples/bwbw.ref.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

// Constants as the reference defines them (mappings.h:4-7).
const double PI_ = 3.14159265358979323846;
const double TWOPI_ = 6.28318530717958647693;
const double TWOPI2_ = 39.47841760435743447534;
const double TWOPI5_ = 9792.62991312900650440772;

// Map ids (phs.h:8-17).
enum { MAP_NO = 0, MAP_SCHANNEL = 1, MAP_TCHANNEL = 2, MAP_UCHANNEL = 3, MAP_RADIATION = 4,
       MAP_COLLINEAR = 5, MAP_INFRARED = 6, MAP_STEP_E = 11, MAP_STEP_H = 12, MAP_ONSHELL = 99 };

// ---------------------------------------------------------------------------
// Forward-mode dual numbers (value + KD tangents).  Only what the algorithm uses.
constexpr int KD = 24;
struct Dual {
  double v;
  double d[KD];
  Dual() : v(0) { std::memset(d, 0, sizeof d); }
  Dual(double x) : v(x) { std::memset(d, 0, sizeof d); }
};
inline Dual operator+(const Dual& a, const Dual& b) { Dual r(a.v + b.v); for (int i = 0; i < KD; i++) r.d[i] = a.d[i] + b.d[i]; return r; }
inline Dual operator-(const Dual& a, const Dual& b) { Dual r(a.v - b.v); for (int i = 0; i < KD; i++) r.d[i] = a.d[i] - b.d[i]; return r; }
inline Dual operator-(const Dual& a) { Dual r(-a.v); for (int i = 0; i < KD; i++) r.d[i] = -a.d[i]; return r; }
inline Dual operator*(const Dual& a, const Dual& b) { Dual r(a.v * b.v); for (int i = 0; i < KD; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i]; return r; }
inline Dual operator/(const Dual& a, const Dual& b) {
  Dual r(a.v / b.v);
  for (int i = 0; i < KD; i++) r.d[i] = (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v);
  return r;
}
inline Dual operator+(const Dual& a, double b) { return a + Dual(b); }
inline Dual operator+(double a, const Dual& b) { return Dual(a) + b; }
inline Dual operator-(const Dual& a, double b) { return a - Dual(b); }
inline Dual operator-(double a, const Dual& b) { return Dual(a) - b; }
inline Dual operator*(const Dual& a, double b) { Dual r(a.v * b); for (int i = 0; i < KD; i++) r.d[i] = a.d[i] * b; return r; }
inline Dual operator*(double a, const Dual& b) { return b * a; }
inline Dual operator/(const Dual& a, double b) { Dual r(a.v / b); for (int i = 0; i < KD; i++) r.d[i] = a.d[i] / b; return r; }
inline Dual operator/(double a, const Dual& b) { return Dual(a) / b; }
inline Dual& operator*=(Dual& a, const Dual& b) { a = a * b; return a; }
inline Dual chain(const Dual& a, double f, double df) { Dual r(f); for (int i = 0; i < KD; i++) r.d[i] = df * a.d[i]; return r; }
inline Dual sqrt(const Dual& a) { double s = std::sqrt(a.v); return chain(a, s, 0.5 / s); }
inline Dual exp(const Dual& a) { double e = std::exp(a.v); return chain(a, e, e); }
inline Dual log(const Dual& a) { return chain(a, std::log(a.v), 1.0 / a.v); }
inline Dual sin(const Dual& a) { return chain(a, std::sin(a.v), std::cos(a.v)); }
inline Dual cos(const Dual& a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
inline Dual tan(const Dual& a) { double t = std::tan(a.v); return chain(a, t, 1.0 + t * t); }
inline Dual atan(const Dual& a) { return chain(a, std::atan(a.v), 1.0 / (1.0 + a.v * a.v)); }
inline double val(const Dual& a) { return a.v; }
inline double val(double a) { return a; }

// Measured sensitivity of the algorithm (used by the parity tests): with a non-zero per-event
// stream (oracle_set_perturbation), every elementary-function result of the double path --
// sqrt, exp, log, sin, cos, tan, atan -- moves by -1, 0 or +1 ulp, drawn from a splitmix64
// sequence.  Two correct FP64 implementations may differ by that much in each of these
// functions, so the spread of the outputs over a few streams measures how far such
// implementations can legitimately drift apart at each event.  Zero, infinite and NaN results
// are left alone (a zero mass stays exactly zero), and so are exact results, which every
// correct implementation returns: a square root whose square is its argument, the root's mass
// sqrt(s) (set_msq), and every function at argument 0.  Constants of fixed beams are evaluated once, outside every stream,
// and never moved (the HIP path takes them from the host as they are); with per-event beams
// they are evaluated inside the event's stream (the HIP path evaluates them on the device).
// With the stream at 0 (the default) every function is the plain libm call, bit for bit.
thread_local uint64_t g_pert = 0;
uint64_t g_pert_seed = 0;
inline uint64_t splitmix64(uint64_t& st) {
  uint64_t z = (st += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double pert(double r) {
  if (!g_pert || r == 0 || !std::isfinite(r)) return r;
  const unsigned k = (unsigned)(splitmix64(g_pert) % 3);
  return k == 0 ? r : std::nextafter(r, k == 1 ? HUGE_VAL : -HUGE_VAL);
}
inline double sqrt(double a) {
  const double r = std::sqrt(a);
  return std::fma(r, r, -a) == 0 ? r : pert(r);
}
inline double sqrt_unperturbed(double a) { return std::sqrt(a); }
inline Dual sqrt_unperturbed(const Dual& a) { return sqrt(a); }
inline double exp(double a) { return a == 0 ? 1.0 : pert(std::exp(a)); }
inline double log(double a) { return pert(std::log(a)); }
inline double sin(double a) { return a == 0 ? a : pert(std::sin(a)); }
inline double cos(double a) { return a == 0 ? 1.0 : pert(std::cos(a)); }
inline double tan(double a) { return a == 0 ? a : pert(std::tan(a)); }
inline double atan(double a) { return a == 0 ? a : pert(std::atan(a)); }

template <class T> T tmax(const T& a, double b) { return val(a) < b ? T(b) : a; }  // std::max(a, b) keeps a on ties
inline double pow2(double q) { return std::pow(q, 2); }  // the reference's pow(q, 2)
inline Dual pow2(const Dual& q) { return q * q; }

// ---------------------------------------------------------------------------
// Process description in the reference's index form: branch index = bincode - 1.
struct Proc {
  int n_in = 2, n_out = 0, n_prt = 0, n_prt_out = 0, root = 0, n_channels = 0;
  std::vector<int> d1, d2, hc;      // [c * n_prt + i]; daughters as indices, -1 = none
  std::vector<int> map_id;          // [c * n_prt_out + i]
  std::vector<double> mass_sum;     // [c * n_prt_out + i]
  std::vector<double> flv;          // [n_in + n_out]
  std::vector<int> fsign;           // [c * n_prt + i] friend sign -1 / 0 / +1, empty: no friends
};

template <class T> struct MapPar {
  std::vector<T> m, w, a, b;        // a, b: [c * n_prt_out * 3 + 3 * i + k]
};

// phs.cu:61-78 (set_mass_sum): minimum invariant mass of every branch = sum of leaf masses.
// The leaf mass index ld2(bincode) + 2 assumes two incoming particles, as the reference does.
void set_mass_sum(const Proc& P, int c, double* ms, int b) {
  if (P.hc[c * P.n_prt + b]) {
    int k1 = P.d1[c * P.n_prt + b], k2 = P.d2[c * P.n_prt + b];
    set_mass_sum(P, c, ms, k1);
    set_mass_sum(P, c, ms, k2);
    ms[b] = ms[k1] + ms[k2];
  } else {
    int ld2 = 0, bb = b + 1;
    while (bb > 1) { bb /= 2; ld2++; }
    ms[b] = P.flv[ld2 + 2];
  }
}

// phs_cpu.cpp:452-531 (init_mapping_constants_cpu): the a[3] (msq map) and b[3] (cos-theta
// map) constants of every (channel, branch), evaluated at one sqrt(s).
template <class T>
void init_constants(const Proc& P, MapPar<T>& M, int c, double sqrts) {
  double m_tot = P.mass_sum[c * P.n_prt_out + P.root];
  for (int i = 0; i < P.n_prt_out; i++) {
    int id = P.map_id[c * P.n_prt_out + i];
    T* a = &M.a[(c * P.n_prt_out + i) * 3];
    T* b = &M.b[(c * P.n_prt_out + i) * 3];
    T m = M.m[c * P.n_prt_out + i], w = M.w[c * P.n_prt_out + i];
    double m_min = P.mass_sum[c * P.n_prt_out + i];
    double m_max = sqrts - m_tot + m_min;
    double msq_min = m_min * m_min, msq_max = m_max * m_max, s = sqrts * sqrts;
    T msq0;
    switch (id) {
      case MAP_NO:
        a[0] = 0; a[1] = msq_max - msq_min; a[2] = a[1] / s; break;
      case MAP_SCHANNEL:
        msq0 = m * m;
        a[0] = atan((msq_min - msq0) / (m * w));
        a[1] = atan((msq_max - msq0) / (m * w));
        a[2] = (a[1] - a[0]) * (m * w) / s; break;
      case MAP_RADIATION: case MAP_COLLINEAR: case MAP_INFRARED:
        msq0 = (id == MAP_RADIATION) ? w * w : m * m;
        a[0] = msq0;
        a[1] = log((msq_max - msq_min) / msq0 + 1);
        a[2] = a[1] / s; break;
      case MAP_TCHANNEL: case MAP_UCHANNEL:
        msq0 = m * m;
        a[0] = msq0;
        a[1] = 2 * log((msq_max - msq_min) / (2 * msq0) + 1);
        a[2] = a[1] / s; break;
      case MAP_STEP_E:
        a[2] = tmax(2 * m * w / (msq_max - msq_min), 0.01);
        a[1] = exp(-(m * m - msq_min) / (msq_max - msq_min) / a[2]);
        a[0] = 1 - a[2] * log((1 + a[1] * exp(1 / a[2])) / (1 + a[1])); break;
      case MAP_STEP_H: {
        a[2] = (m * m - msq_min) / (msq_max - msq_min);
        a[1] = tmax(pow2(2 * m * w / (msq_max - msq_min)) / a[2], 0.000001);
        a[0] = (1 + sqrt(1 + 4 * a[1] / (1 - a[2]))) / 2; break;
      }
      default: break;
    }
    switch (id) {
      case MAP_TCHANNEL: case MAP_UCHANNEL: case MAP_COLLINEAR:
        b[0] = m * m / s; b[1] = log((b[0] + 1) / b[0]); b[2] = 0; break;
      default:
        b[0] = 0; b[1] = 0; b[2] = 0; break;
    }
  }
}

// phs_cpu.cpp:12-61 (mapping_msq_from_x_cpu).
//
// kappa (may be NULL) receives the condition number of the result: for every subtraction
// a - b on the way, (|a| + |b|) / |a - b|, and for tan near +-pi/2 its own amplification.
template <class T>
void map_msq(int type, const T& x, double s, const T& m, const T& w, double msq_min, double msq_max,
             const T* a, T* msq, T* factor, double* kappa = nullptr) {
  T msq1, tmp, z;
  double kap = 1;
  auto cancel = [](double u, double v) {  // condition of u - v
    double d = std::fabs(u - v);
    return d > 0 ? (std::fabs(u) + std::fabs(v)) / d : 1e300;
  };
  switch (type) {
    case MAP_NO:
      *msq = (1 - x) * msq_min + x * msq_max;
      *factor = T((msq_max - msq_min) / s);
      break;
    case MAP_SCHANNEL:
      z = (1 - x) * a[0] + x * a[1];
      if (-PI_ / 2 < val(z) && val(z) < PI_ / 2) {
        tmp = tan(z);
        *msq = m * (m + w * tmp);
        *factor = a[2] * (1 + tmp * tmp);
        {
          double t = val(tmp), zz = std::fabs(val(z)) + std::fabs(val(a[0])) + std::fabs(val(a[1]));
          double mm = std::fabs(val(m)), ww = std::fabs(val(w));
          kap = std::max(kap, mm * ww * (1 + t * t) * zz / std::max(std::fabs(val(*msq)), 1e-300));
          kap = std::max(kap, cancel(-mm * mm, ww * mm * t));
        }
      } else {
        *msq = 0; *factor = 0;
      }
      break;
    case MAP_COLLINEAR: case MAP_INFRARED: case MAP_RADIATION:
      msq1 = a[0] * exp(x * a[1]);
      *msq = msq1 - a[0] + msq_min;
      *factor = a[2] * msq1;
      kap = (std::fabs(val(msq1)) + std::fabs(val(a[0])) + msq_min) / std::max(std::fabs(val(*msq)), 1e-300);
      break;
    case MAP_TCHANNEL: case MAP_UCHANNEL:
      if (val(x) < 0.5) {
        msq1 = a[0] * exp(x * a[1]);
        *msq = msq1 - a[0] + msq_min;
        kap = (std::fabs(val(msq1)) + std::fabs(val(a[0])) + msq_min) / std::max(std::fabs(val(*msq)), 1e-300);
      } else {
        msq1 = a[0] * exp((1 - x) * a[1]);
        *msq = -(msq1 - a[0]) + msq_max;
        kap = (std::fabs(val(msq1)) + std::fabs(val(a[0])) + msq_max) / std::max(std::fabs(val(*msq)), 1e-300);
      }
      *factor = a[2] * msq1;
      break;
    case MAP_STEP_H:
      z = a[1] / (a[0] - x) - a[1] / a[0] + a[2] * x;
      *msq = z * msq_max + (1 - z) * msq_min;
      *factor = (a[1] / ((a[0] - x) * (a[0] - x)) + a[2]) * (msq_max - msq_min) / s;
      kap = cancel(val(a[1] / (a[0] - x)), val(a[1] / a[0] - a[2] * x)) * std::fabs(val(z)) * (msq_max + msq_min) /
            std::max(std::fabs(val(*msq)), 1e-300);
      break;
    case MAP_STEP_E:
      tmp = exp(-x * a[0] / a[2]) * (1 + a[1]);
      z = -a[2] * log(tmp - a[1]);
      *msq = z * msq_max + (1 - z) * msq_min;
      *factor = a[0] / (1 - a[1] / tmp) * (msq_max - msq_min) / s;
      kap = cancel(val(tmp), val(a[1])) * std::max(1.0, std::fabs(val(a[2])) * (msq_max + msq_min) /
                                                          std::max(std::fabs(val(*msq)), 1e-300));
      break;
  }
  if (kappa) *kappa = kap;
}

// phs_cpu.cpp:112-146 (mapping_ct_from_x_cpu).
template <class T>
void map_ct(int type, const T& x, const T* b, T* ct, T* st, T* factor) {
  T tmp;
  switch (type) {
    case MAP_NO: case MAP_SCHANNEL: case MAP_INFRARED: case MAP_RADIATION: case MAP_STEP_E: case MAP_STEP_H:
      tmp = 2 * (1 - x);
      *ct = 1 - tmp;
      *st = sqrt(tmp * (2 - tmp));
      *factor = 1;
      break;
    case MAP_COLLINEAR: case MAP_TCHANNEL: case MAP_UCHANNEL:
      if (val(x) < 0.5) {
        tmp = b[0] * exp(2 * x * b[1]);
        *ct = tmp - b[0] - 1;
      } else {
        tmp = b[0] * exp(2 * (1 - x) * b[1]);
        *ct = -(tmp - b[0]) + 1;
      }
      if (val(*ct) >= -1 && val(*ct) <= 1) {
        *st = sqrt(1 - *ct * *ct);
        *factor = tmp * b[1];
      } else {
        *ct = 1; *st = 0; *factor = 0;
      }
      break;
  }
}

template <class T> struct Ev {
  std::vector<T> p2;        // msq per branch index (prt[].p2 of the reference)
  std::vector<T> pd;        // p_decay per branch index
  std::vector<T> mom;       // 4 per branch index (prt[].p)
  std::vector<double> scale;  // 4 per branch index: forward-error scale of mom
  double cond = 1;            // condition number of the event (see set_msq, set_angles)
  double s_friend = 0;        // s of the cos-theta maps of branches with a friend: 4 E1 E2
};

// phs_cpu.cpp:174-234 (set_msq_cpu): post-order recursion, msq randoms consumed in
// post-order (k1 subtree, k2 subtree, self), root excluded; early return on !ok.
template <class T>
void set_msq(const Proc& P, const MapPar<T>& M, int c, int b, double m_tot, const T* x, int* id_x,
             double sqrts, Ev<T>& E, T* factor, T* volume, bool* ok) {
  int k1 = P.d1[c * P.n_prt + b], k2 = P.d2[c * P.n_prt + b];
  T f1, f2, v1, v2;
  if (P.hc[c * P.n_prt + k1]) {
    set_msq(P, M, c, k1, m_tot, x, id_x, sqrts, E, &f1, &v1, ok);
    if (!*ok) return;
  } else { f1 = 1; v1 = 1; }
  if (P.hc[c * P.n_prt + k2]) {
    set_msq(P, M, c, k2, m_tot, x, id_x, sqrts, E, &f2, &v2, ok);
    if (!*ok) return;
  } else { f2 = 1; v2 = 1; }

  double m_max;
  if (b == P.root) {
    E.p2[b] = sqrts * sqrts;
    m_max = sqrts;
    *factor = f1 * f2;
    *volume = v1 * v2 / (4 * TWOPI5_);
  } else {
    int ib = c * P.n_prt_out + b;
    double m_min = P.mass_sum[ib];
    m_max = sqrts - m_tot + m_min;
    double kap = 1;
    map_msq(P.map_id[ib], x[*id_x], sqrts * sqrts, M.m[ib], M.w[ib], m_min * m_min, m_max * m_max,
            &M.a[ib * 3], &E.p2[b], factor, &kap);
    E.cond = std::max(E.cond, kap);
    (*id_x)++;
    // The reference tests a local that is always 0 here (phs_cpu.cpp:202,210), so this
    // branch is always taken.
    *factor *= f1 * f2;
    *volume = v1 * v2 * sqrts * sqrts / (4 * TWOPI2_);
  }
  if (*ok) {
    T this_msq = E.p2[b], msq1 = E.p2[k1], msq2 = E.p2[k2];
    // the root's mass: sqrt(s) of s = sqrt(s)^2, which a correctly rounded square root returns
    // exactly and the cut m <= sqrt(s) below relies on; it is never perturbed (see pert())
    T m1 = sqrt(msq1), m2 = sqrt(msq2), m = b == P.root ? sqrt_unperturbed(this_msq) : sqrt(this_msq);
    T lda = (this_msq - msq1 - msq2) * (this_msq - msq1 - msq2) - 4 * msq1 * msq2;
    // Condition number of lambda: the terms it is formed from over lambda itself.  Near a
    // threshold (lambda -> 0) an ulp of difference in any msq moves sqrt(lambda), the factor
    // and the decay momenta by ~eps * cond.
    if (val(lda) > 0) {
      double t = std::fabs(val(this_msq)) + std::fabs(val(msq1)) + std::fabs(val(msq2));
      E.cond = std::max(E.cond, t * t / val(lda));
    }
    if (val(lda) > 0 && val(m) > val(m1 + m2) && val(m) <= m_max) {
      E.pd[k1] = sqrt(lda) / (2 * m);
      E.pd[k2] = -E.pd[k1];
      *factor *= sqrt(lda) / this_msq;
    } else {
      *ok = false;
    }
  }
}

// The boost of a branch whose polar axis is an incoming beam, its "friend"
// (phs.cu:853-910, GPU-only in the reference).  The beam k = (1, 0, 0, f) (the reference
// scales it by the beam energy, phs.cu:885-886, which the normalisation removes) seen from
// the branch's rest frame is n = spatial part of B_z(bg)^-1 L0^-1 k, L0^-1 = eta L0^T eta
// (phs.cu:888-890, whose n_z has the opposite sign on the bg L0[3][0] term).  R is the
// rotation about e_z x n that takes e_z to n / |n| (Rodrigues' formula; for n along -e_z the
// half turn about e_x), where the reference's L_rot_to_2nd (phs.cu:753-781) builds a matrix
// that is not orthogonal.  L1 = B_z(bg) (1 (+) R R_z(phi) R_y(theta)), phs.cu:896-906.
// cond grows by the conditioning of the axis: the terms n is summed from over |n|, times the
// sensitivity (1 - n_z) / |n_xy| of the rotation axis near n = -e_z.
template <class T>
void friend_l1(int fs, const T L0[4][4], const T& bg, const T& gamma, const T& ct, const T& st, const T& cp,
               const T& sp, T L1[4][4], double* cond) {
  const double f = fs;
  T u0 = L0[0][0] - f * L0[3][0];
  T n0 = -L0[0][1] + f * L0[3][1];
  T n1 = -L0[0][2] + f * L0[3][2];
  T u3 = -L0[0][3] + f * L0[3][3];
  T n2 = -bg * u0 + gamma * u3;
  T nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
  T a0 = n0 / nn, a1 = n1 / nn, cc = n2 / nn;
  T sa = sqrt(a0 * a0 + a1 * a1);
  T kx = val(sa) > 0 ? T(-a1 / sa) : T(1.0);
  T ky = val(sa) > 0 ? T(a0 / sa) : T(0.0);
  T h = 1 - cc;
  T R[3][3] = {{cc + h * kx * kx, h * kx * ky, a0}, {h * kx * ky, cc + h * ky * ky, a1}, {-a0, -a1, cc}};
  T Rot[3][3] = {{cp * ct, -sp, cp * st}, {sp * ct, cp, sp * st}, {-st, T(0.0), ct}};
  T Q[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      // the kernel skips the structural zero Rot[2][1]; adding 0 changes no value
      Q[i][j] = R[i][0] * Rot[0][j] + R[i][1] * Rot[1][j];
      if (j != 1) Q[i][j] = Q[i][j] + R[i][2] * Rot[2][j];
    }
  L1[0][0] = gamma; L1[0][1] = bg * Q[2][0];    L1[0][2] = bg * Q[2][1];    L1[0][3] = bg * Q[2][2];
  L1[1][0] = 0;     L1[1][1] = Q[0][0];         L1[1][2] = Q[0][1];         L1[1][3] = Q[0][2];
  L1[2][0] = 0;     L1[2][1] = Q[1][0];         L1[2][2] = Q[1][1];         L1[2][3] = Q[1][2];
  L1[3][0] = bg;    L1[3][1] = gamma * Q[2][0]; L1[3][2] = gamma * Q[2][1]; L1[3][3] = gamma * Q[2][2];
  const double ab = std::fabs(val(bg)), g = std::fabs(val(gamma));
  double t = 0;
  for (int j = 1; j < 3; j++) t += std::fabs(val(L0[0][j])) + std::fabs(val(L0[3][j]));
  t += ab * (std::fabs(val(L0[0][0])) + std::fabs(val(L0[3][0]))) + g * (std::fabs(val(L0[0][3])) + std::fabs(val(L0[3][3])));
  double k = t / std::max(val(nn), 1e-300);
  if (val(sa) > 0) k *= std::max(1.0, val(h) / val(sa));
  *cond = std::max(*cond, k);
}

// phs_cpu.cpp:283-345 (set_angles_cpu): pre-order recursion; phi then cos-theta randoms
// per internal branch; L_new = L0 * B_z(bg) R_z(phi) R_y(theta).
//
// Alongside, the componentwise forward-error scale of every momentum component is carried:
// A = |L1_root| |L1_a1| ... |L1_parent| composed in absolute values, and
// scale^mu = A[mu][0] |E| + A[mu][3] |p| bounds the magnitude of the terms that form p^mu
// (E.scale, per branch).  Two correct double-precision evaluations of the chain can differ
// by a small multiple of eps * scale, however small p^mu itself is.
template <class T>
void set_angles(const Proc& P, const MapPar<T>& M, int c, int b, const T* x, int* idx_x, T* factor,
                Ev<T>& E, const T L0[4][4], const double A0[4][4]) {
  T p = E.pd[b];
  T m = sqrt(E.p2[b]);
  T En = sqrt(E.p2[b] + p * p);
  for (int i = 0; i < 4; i++) E.mom[4 * b + i] = L0[i][0] * En + L0[i][3] * p;
  for (int i = 0; i < 4; i++) E.scale[4 * b + i] = A0[i][0] * std::fabs(val(En)) + A0[i][3] * std::fabs(val(p));
  if (P.hc[c * P.n_prt + b]) {
    int k1 = P.d1[c * P.n_prt + b], k2 = P.d2[c * P.n_prt + b];
    T bg = val(m) > 0 ? p / m : T(0);
    T gamma = sqrt(1 + bg * bg);
    T phi = x[*idx_x] * TWOPI_;
    (*idx_x)++;
    T cp = cos(phi), sp = sin(phi);
    T xx = x[*idx_x];
    (*idx_x)++;
    int ib = c * P.n_prt_out + b;
    T ct, st, f;
    const int fs = P.fsign.empty() ? 0 : P.fsign[c * P.n_prt + b];
    if (fs != 0) {
      // a branch with a friend evaluates its cos-theta map at s = 4 E1 E2 (phs.cu:879;
      // b[] as mappings_gpu.hpp:186-188)
      T bf[3];
      bf[0] = M.m[ib] * M.m[ib] / E.s_friend;
      bf[1] = log((bf[0] + 1) / bf[0]);
      bf[2] = 0;
      map_ct(P.map_id[ib], xx, bf, &ct, &st, &f);
    } else {
      map_ct(P.map_id[ib], xx, &M.b[ib * 3], &ct, &st, &f);
    }
    *factor *= f;
    // sin theta = sqrt(1 - ct^2) of the peaked map loses accuracy as 1 / sin theta, and
    // ct = tmp - b0 - 1 cancels near ct = 0
    if (val(st) > 0 && P.map_id[ib] != MAP_NO && val(f) != 1.0) {
      E.cond = std::max(E.cond, 1.0 / std::fabs(val(st)));
      double c = std::fabs(val(ct));
      if (c > 0) E.cond = std::max(E.cond, (1 + std::fabs(val(M.b[ib * 3])) + 1) / c);
    }
    T L1[4][4];
    if (fs == 0) {
      L1[0][0] = gamma; L1[0][1] = -bg * st;    L1[0][2] = 0;  L1[0][3] = bg * ct;
      L1[1][0] = 0;     L1[1][1] = ct * cp;     L1[1][2] = -sp; L1[1][3] = st * cp;
      L1[2][0] = 0;     L1[2][1] = ct * sp;     L1[2][2] = cp; L1[2][3] = st * sp;
      L1[3][0] = bg;    L1[3][1] = -gamma * st; L1[3][2] = 0;  L1[3][3] = gamma * ct;
    } else {
      friend_l1(fs, L0, bg, gamma, ct, st, cp, sp, L1, &E.cond);
    }
    T Ln[4][4];
    double An[4][4];
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) {
        Ln[i][j] = 0;
        An[i][j] = 0;
        for (int k = 0; k < 4; k++) {
          Ln[i][j] = Ln[i][j] + L0[i][k] * L1[k][j];
          An[i][j] += A0[i][k] * std::fabs(val(L1[k][j]));
        }
      }
    set_angles(P, M, c, k1, x, idx_x, factor, E, Ln, An);
    set_angles(P, M, c, k2, x, idx_x, factor, E, Ln, An);
  }
}

// Import the caller's tree in the reference's bincode form (whizard_gpu_phs.cpp:79-87,
// daughters as bincodes, 0 = none) and convert to indices as init_phs_gpu does
// (phs.cu:392-397).
bool build_proc(Proc& P, int n_in, int n_out, int n_channels, const int* d1, const int* d2,
                const int* hc, const int* map_ids, const double* flv) {
  if (n_in != 2 || n_out < 2 || n_out > 12 || n_channels < 1) return false;
  P.n_in = n_in; P.n_out = n_out; P.n_channels = n_channels;
  P.n_prt = (1 << (n_in + n_out)) - 1;
  P.n_prt_out = (1 << n_out) - 1;
  P.root = P.n_prt_out - 1;
  P.d1.assign(d1, d1 + (size_t)n_channels * P.n_prt);
  P.d2.assign(d2, d2 + (size_t)n_channels * P.n_prt);
  for (auto& v : P.d1) v -= 1;
  for (auto& v : P.d2) v -= 1;
  P.hc.assign(hc, hc + (size_t)n_channels * P.n_prt);
  P.map_id.assign(map_ids, map_ids + (size_t)n_channels * P.n_prt_out);
  P.flv.assign(flv, flv + n_in + n_out);
  P.mass_sum.assign((size_t)n_channels * P.n_prt_out, 0.0);
  for (int c = 0; c < n_channels; c++) set_mass_sum(P, c, &P.mass_sum[(size_t)c * P.n_prt_out], P.root);
  return true;
}

// Friend signs from the caller's friends array [n_channels][n_prt] of incoming bincodes
// (whizard_gpu_phs.cpp:79-87), as extract_boost_origins reads it (phs.cu:276-284):
// 2^n_out gives -1, 2^(n_out + 1) gives +1, anything else no friend.
void set_friends(Proc& P, const int* friends) {
  P.fsign.assign((size_t)P.n_channels * P.n_prt, 0);
  if (!friends) return;
  const int in1 = 1 << P.n_out;
  for (size_t i = 0; i < P.fsign.size(); i++)
    P.fsign[i] = friends[i] == in1 ? -1 : friends[i] == 2 * in1 ? 1 : 0;
}

// One event, phs_cpu.cpp:582-600 (reset prt, init_msq_cpu, set_msq_cpu, set_angles_cpu),
// then the lab-frame boost of the final leaves, phs.cu:1286-1301, when E1 != E2.
template <class T>
bool one_event(const Proc& P, const MapPar<T>& M, int c, double sqrts, double bg_lab, double s_friend, const T* x,
               Ev<T>& E, T* factor, T* volume, T* p_out, double* msq_out, double* scale_out, double* cond_out) {
  int n = P.n_prt;
  E.p2.assign(n, T(0)); E.pd.assign(n, T(0)); E.mom.assign(4 * n, T(0)); E.scale.assign(4 * n, 0.0);
  E.cond = 1;
  E.s_friend = s_friend;
  // phs_cpu.cpp:541-547 (init_msq_cpu): leaf msq = flavour mass squared.
  for (int i = 0, p = 1; i < P.n_out; i++, p *= 2) E.p2[p - 1] = P.flv[P.n_in + i] * P.flv[P.n_in + i];
  bool ok = true;
  int id_x = 0;
  double m_tot = P.mass_sum[c * P.n_prt_out + P.root];
  set_msq(P, M, c, P.root, m_tot, x, &id_x, sqrts, E, factor, volume, &ok);
  if (msq_out)
    for (int i = 0; i < P.n_prt_out; i++) msq_out[i] = val(E.p2[i]);
  if (ok) {
    T L0[4][4];
    double A0[4][4];
    for (int i = 0; i < 4; i++)
      for (int j = 0; j < 4; j++) { L0[i][j] = (i == j) ? 1.0 : 0.0; A0[i][j] = (i == j) ? 1.0 : 0.0; }
    set_angles(P, M, c, P.root, x, &id_x, factor, E, L0, A0);
    for (int j = 0, b = 1; j < P.n_out; j++, b *= 2)
      for (int mu = 0; mu < 4; mu++) {
        p_out[4 * j + mu] = E.mom[4 * (b - 1) + mu];
        if (scale_out) scale_out[4 * j + mu] = E.scale[4 * (b - 1) + mu];
      }
    if (bg_lab != 0) {
      double g = std::sqrt(1 + bg_lab * bg_lab);
      for (int j = 0; j < P.n_out; j++) {
        T p0 = g * p_out[4 * j] + bg_lab * p_out[4 * j + 3];
        T p3 = bg_lab * p_out[4 * j] + g * p_out[4 * j + 3];
        p_out[4 * j] = p0; p_out[4 * j + 3] = p3;
        if (scale_out) {
          double s0 = g * scale_out[4 * j] + std::fabs(bg_lab) * scale_out[4 * j + 3];
          double s3 = std::fabs(bg_lab) * scale_out[4 * j] + g * scale_out[4 * j + 3];
          scale_out[4 * j] = s0; scale_out[4 * j + 3] = s3;
        }
      }
    }
  }
  if (cond_out) *cond_out = E.cond;
  return ok;
}

// ---------------------------------------------------------------------------
// Inverse mapping: the randoms and the phase-space factor under which channel c would have
// produced a given event (phs.cu:626-675, 924-1284 on the GPU, driven per channel by
// phs.cu:1499-1554; CPU counterpart phs_cpu.cpp:63-110, 148-172, 236-281, 347-449).  Follows
// the GPU reference's steps with the defects of both versions corrected, so that the inverse
// of the channel that generated an event returns its randoms and factor:
//   * STEP_H: the root of the quadratic with a1 / (a0 a2) (phs_cpu.cpp:97-99), where the GPU
//     puts z1 (mappings_gpu.hpp:167-169);
//   * T/U channel, upper half: x = 1 - log(z / a0) / a1 (mappings_gpu.hpp:119-121), where the
//     CPU drops the "1 -" (phs_cpu.cpp:93);
//   * phi = atan2(n_y, n_x) in [0, 2 pi) (phs.cu:912-915), where the CPU takes atan(n_y / n_x);
//   * the angles of a branch are those of its first daughter (phs.cu:938-962), where the CPU
//     uses the branch's own momentum;
//   * lambda <= 0 gives factor 0 and p_decay 0 (phs_cpu.cpp:270-279) instead of NaN;
//   * cos theta = n_z / |n| and sin theta = |n_xy| / |n| (phs.cu:917-920 takes
//     sqrt(1 - cos^2)), and cos phi, sin phi = n_x, n_y over |n_xy| instead of cos(atan2)
//     (phs.cu:1084-1085);
//   * a branch with a friend measures its angles from the friend beam as the forward pass
//     does (friend_l1), not with the reference's non-orthogonal rotation (phs.cu:1128-1240).
// Inputs are the event's momenta in the centre-of-mass frame.

// x from msq, mappings_gpu.hpp:15-172 (step_h: phs_cpu.cpp:95-101).
void map_x_from_msq(int type, double msq, double s, double m, double w, double msq_min, double msq_max,
                    const double* a, double* x, double* factor) {
  double z, tmp;
  switch (type) {
    case MAP_NO:
      *x = (msq - msq_min) / (msq_max - msq_min);
      *factor = (msq_max - msq_min) / s;
      break;
    case MAP_SCHANNEL:
      z = (msq - m * m) / (m * w);
      *x = (std::atan(z) - a[0]) / (a[1] - a[0]);
      *factor = a[2] * (1 + z * z);
      break;
    case MAP_COLLINEAR: case MAP_INFRARED: case MAP_RADIATION:
      z = msq - msq_min + a[0];
      *x = std::log(z / a[0]) / a[1];
      *factor = a[2] * z;
      break;
    case MAP_TCHANNEL: case MAP_UCHANNEL:
      if (msq < (msq_max + msq_min) / 2) {
        z = msq - msq_min + a[0];
        *x = std::log(z / a[0]) / a[1];
      } else {
        z = msq_max - msq + a[0];
        *x = 1 - std::log(z / a[0]) / a[1];
      }
      *factor = a[2] * z;
      break;
    case MAP_STEP_E:
      z = (msq - msq_min) / (msq_max - msq_min);
      tmp = 1 + a[1] * std::exp(z / a[2]);
      *x = (z - a[2] * std::log(tmp / (1 + a[1]))) / a[0];
      *factor = a[0] * tmp * (msq_max - msq_min) / s;
      break;
    case MAP_STEP_H:
      z = (msq - msq_min) / (msq_max - msq_min);
      tmp = a[1] / (a[0] * a[2]);
      *x = ((a[0] + z / a[2] + tmp) -
            std::sqrt((a[0] - z / a[2]) * (a[0] - z / a[2]) + 2 * tmp * (a[0] + z / a[2]) + tmp * tmp)) / 2;
      *factor = (a[1] / (a[0] - *x) / (a[0] - *x) + a[2]) * (msq_max - msq_min) / s;
      break;
    default:
      *x = 0; *factor = 0;
      break;
  }
}

// x from cos theta, mappings_gpu.hpp:181-219 (phs_cpu.cpp:148-172).
void map_x_from_ct(int type, double ct, const double* b, double* x, double* factor) {
  switch (type) {
    case MAP_COLLINEAR: case MAP_TCHANNEL: case MAP_UCHANNEL: {
      double t;
      if (ct < 0) {
        t = ct + b[0] + 1;
        *x = std::log(t / b[0]) / (2 * b[1]);
      } else {
        t = -ct + b[0] + 1;
        *x = 1 - std::log(t / b[0]) / (2 * b[1]);
      }
      *factor = t * b[1];
      break;
    }
    default:
      *x = (ct + 1) / 2;
      *factor = 1;
      break;
  }
}

struct InvState {
  std::vector<double> P, msq, pd;   // per branch index: four-momentum, msq, p_decay
};

// phs.cu:1262-1284 (_combine_particles): branch momenta as sums of their daughters, msq of
// every branch from its four-momentum (leaves included).
void combine(const Proc& P, int c, int b, InvState& S) {
  if (!P.hc[c * P.n_prt + b]) return;
  int k1 = P.d1[c * P.n_prt + b], k2 = P.d2[c * P.n_prt + b];
  combine(P, c, k1, S);
  combine(P, c, k2, S);
  for (int i = 0; i < 4; i++) S.P[4 * b + i] = S.P[4 * k1 + i] + S.P[4 * k2 + i];
}

// phs.cu:626-675 (_apply_msq_inv), post-order as the forward pass draws the msq randoms.
void inv_msq(const Proc& P, const MapPar<double>& M, int c, int b, double sqrts, double m_tot, InvState& S,
             double* x, int* id_x, double* factor) {
  int k1 = P.d1[c * P.n_prt + b], k2 = P.d2[c * P.n_prt + b];
  double f1 = 1, f2 = 1;
  if (P.hc[c * P.n_prt + k1]) inv_msq(P, M, c, k1, sqrts, m_tot, S, x, id_x, &f1);
  if (P.hc[c * P.n_prt + k2]) inv_msq(P, M, c, k2, sqrts, m_tot, S, x, id_x, &f2);
  int ib = c * P.n_prt_out + b;
  if (b == P.root) {
    S.msq[b] = sqrts * sqrts;
    *factor = f1 * f2;
  } else {
    double m_min = P.mass_sum[ib], m_max = sqrts - m_tot + m_min;
    double f;
    map_x_from_msq(P.map_id[ib], S.msq[b], sqrts * sqrts, M.m[ib], M.w[ib], m_min * m_min, m_max * m_max,
                   &M.a[3 * ib], &x[*id_x], &f);
    (*id_x)++;
    *factor = f * f1 * f2;
  }
  double msq0 = S.msq[b], msq1 = S.msq[k1], msq2 = S.msq[k2];
  double lda = (msq0 - msq1 - msq2) * (msq0 - msq1 - msq2) - 4 * msq1 * msq2;
  if (lda > 0) {
    double m0 = std::sqrt(msq0);
    S.pd[k1] = std::sqrt(lda) / (2 * m0);
    S.pd[k2] = -S.pd[k1];
    *factor *= std::sqrt(lda) / msq0;
  } else {
    S.pd[k1] = S.pd[k2] = 0;
    *factor = 0;
  }
}

// phs.cu:924-1241 (_create_boosts_inv_*), pre-order.  G maps centre-of-mass vectors into the
// frame of the parent's boost matrix (G = L_parent^-1); the branch's randoms are the angles
// of its first daughter in B_z(-bg) G.
void inv_angles(const Proc& P, const MapPar<double>& M, int c, int b, double s_friend, InvState& S,
                const double G[4][4], double* x, int* id_x, double* factor) {
  if (!P.hc[c * P.n_prt + b]) return;
  int k1 = P.d1[c * P.n_prt + b], k2 = P.d2[c * P.n_prt + b];
  int ib = c * P.n_prt_out + b;
  double m = std::sqrt(S.msq[b]);
  double bg = m > 0 ? S.pd[b] / m : 0;
  double g = std::sqrt(1 + bg * bg);
  double W[4][4];
  for (int j = 0; j < 4; j++) {
    W[0][j] = g * G[0][j] - bg * G[3][j];
    W[1][j] = G[1][j];
    W[2][j] = G[2][j];
    W[3][j] = -bg * G[0][j] + g * G[3][j];
  }
  const int fs = P.fsign.empty() ? 0 : P.fsign[c * P.n_prt + b];
  double v[3];
  for (int i = 0; i < 3; i++)
    v[i] = W[i + 1][0] * S.P[4 * k1] + W[i + 1][1] * S.P[4 * k1 + 1] + W[i + 1][2] * S.P[4 * k1 + 2] +
           W[i + 1][3] * S.P[4 * k1 + 3];
  double Rt[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};   // R^T of the friend rotation
  if (fs != 0) {
    // the beam k = (1, 0, 0, f) in the same frame; R takes e_z to it (friend_l1)
    double n0 = W[1][0] + fs * W[1][3], n1 = W[2][0] + fs * W[2][3], n2 = W[3][0] + fs * W[3][3];
    double nn = std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
    double a0 = n0 / nn, a1 = n1 / nn, cc = n2 / nn;
    double sa = std::sqrt(a0 * a0 + a1 * a1);
    double kx = sa > 0 ? -a1 / sa : 1.0, ky = sa > 0 ? a0 / sa : 0.0, h = 1 - cc;
    double R[3][3] = {{cc + h * kx * kx, h * kx * ky, a0}, {h * kx * ky, cc + h * ky * ky, a1}, {-a0, -a1, cc}};
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) Rt[i][j] = R[j][i];
    double u[3] = {v[0], v[1], v[2]};
    for (int i = 0; i < 3; i++) v[i] = Rt[i][0] * u[0] + Rt[i][1] * u[1] + Rt[i][2] * u[2];
  }
  double rho = std::sqrt(v[0] * v[0] + v[1] * v[1]);
  double vv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  double phi = std::atan2(v[1], v[0]);
  if (phi < 0) phi += TWOPI_;
  double cp = rho > 0 ? v[0] / rho : 1.0, sp = rho > 0 ? v[1] / rho : 0.0;
  double ct = v[2] / vv, st = rho / vv;
  x[*id_x] = phi / TWOPI_;
  (*id_x)++;
  double bb[3] = {M.b[3 * ib], M.b[3 * ib + 1], M.b[3 * ib + 2]};
  if (fs != 0) {
    bb[0] = M.m[ib] * M.m[ib] / s_friend;
    bb[1] = std::log((bb[0] + 1) / bb[0]);
    bb[2] = 0;
  }
  double f;
  map_x_from_ct(P.map_id[ib], ct, bb, &x[*id_x], &f);
  (*id_x)++;
  *factor *= f;
  // G_b = R_y(-theta) R_z(-phi) R^T W  (= L_b^-1)
  double Gn[4][4];
  for (int j = 0; j < 4; j++) {
    double w1 = Rt[0][0] * W[1][j] + Rt[0][1] * W[2][j] + Rt[0][2] * W[3][j];
    double w2 = Rt[1][0] * W[1][j] + Rt[1][1] * W[2][j] + Rt[1][2] * W[3][j];
    double w3 = Rt[2][0] * W[1][j] + Rt[2][1] * W[2][j] + Rt[2][2] * W[3][j];
    double t1 = cp * w1 + sp * w2;
    Gn[0][j] = W[0][j];
    Gn[1][j] = ct * t1 - st * w3;
    Gn[2][j] = -sp * w1 + cp * w2;
    Gn[3][j] = st * t1 + ct * w3;
  }
  inv_angles(P, M, c, k1, s_friend, S, Gn, x, id_x, factor);
  inv_angles(P, M, c, k2, s_friend, S, Gn, x, id_x, factor);
}

}  // namespace

extern "C" {

// Sizes of the reference's derived quantities (main.cpp:271-285).
int oracle_n_prt(int n_in, int n_out) { return (1 << (n_in + n_out)) - 1; }
int oracle_n_prt_out(int n_out) { return (1 << n_out) - 1; }

// Mass sums per (channel, branch index), phs.cu:61-78.  Returns 0 on success.
int oracle_mass_sums(int n_in, int n_out, int n_channels, const int* d1, const int* d2, const int* hc,
                     const int* map_ids, const double* flv, double* out) {
  Proc P;
  if (!build_proc(P, n_in, n_out, n_channels, d1, d2, hc, map_ids, flv)) return 1;
  std::memcpy(out, P.mass_sum.data(), P.mass_sum.size() * sizeof(double));
  return 0;
}

// Mapping constants a[3], b[3] per (channel, branch index) at one sqrt(s),
// phs_cpu.cpp:452-531.  a_out, b_out: [n_channels][n_prt_out][3].
int oracle_constants(int n_in, int n_out, int n_channels, const int* d1, const int* d2, const int* hc,
                     const int* map_ids, const double* map_m, const double* map_w, const double* flv,
                     double sqrts, double* a_out, double* b_out) {
  Proc P;
  if (!build_proc(P, n_in, n_out, n_channels, d1, d2, hc, map_ids, flv)) return 1;
  MapPar<double> M;
  size_t nb = (size_t)n_channels * P.n_prt_out;
  M.m.assign(map_m, map_m + nb); M.w.assign(map_w, map_w + nb);
  M.a.assign(3 * nb, 0.0); M.b.assign(3 * nb, 0.0);
  for (int c = 0; c < n_channels; c++) init_constants(P, M, c, sqrts);
  std::memcpy(a_out, M.a.data(), 3 * nb * sizeof(double));
  std::memcpy(b_out, M.b.data(), 3 * nb * sizeof(double));
  return 0;
}

// Single-map probes for closed-form tests: phs_cpu.cpp:12-61 and 112-146.
void oracle_map_msq(int type, double x, double s, double m, double w, double msq_min, double msq_max,
                    const double* a, double* msq, double* factor) {
  map_msq<double>(type, x, s, m, w, msq_min, msq_max, a, msq, factor);
}
void oracle_map_ct(int type, double x, const double* b, double* ct, double* st, double* factor) {
  map_ct<double>(type, x, b, ct, st, factor);
}

// Forward generation of n_events.
//   trees: d1, d2, hc [n_channels][n_prt] (bincodes, 0 = none), map_ids / map_m / map_w
//   [n_channels][n_prt_out] indexed by bincode - 1, flv [n_in + n_out] (incoming first).
//   E1, E2: per-event beam energies, or NULL for both beams at sqrts / 2.
//   channels [n_events] 0-based; x [n_events][n_x] event-major.
//   Outputs: p [n_events][n_out][4], factors, volumes [n_events], oks [n_events] (0/1),
//   msq [n_events][n_prt_out] (may be NULL), scale [n_events][n_out][4] the forward-error
//   scale of every momentum component (may be NULL; see set_angles), cond [n_events] the
//   event's condition number (may be NULL; see set_msq).  Non-ok events keep p = 0 and whatever factor
//   and volume the early return left, as the reference does.  Returns 0 on success.
int oracle_forward_friends(int n_in, int n_out, int n_channels, const int* d1, const int* d2, const int* hc,
                           const int* friends, const int* map_ids, const double* map_m, const double* map_w,
                           const double* flv, double sqrts, const double* E1, const double* E2, long n_events,
                           const int* channels, int n_x, const double* x, double* p, double* factors,
                           double* volumes, unsigned char* oks, double* msq, double* scale, double* cond);

int oracle_forward(int n_in, int n_out, int n_channels, const int* d1, const int* d2, const int* hc,
                   const int* map_ids, const double* map_m, const double* map_w, const double* flv,
                   double sqrts, const double* E1, const double* E2, long n_events, const int* channels,
                   int n_x, const double* x, double* p, double* factors, double* volumes,
                   unsigned char* oks, double* msq, double* scale, double* cond) {
  return oracle_forward_friends(n_in, n_out, n_channels, d1, d2, hc, nullptr, map_ids, map_m, map_w, flv, sqrts,
                                E1, E2, n_events, channels, n_x, x, p, factors, volumes, oks, msq, scale, cond);
}

// oracle_forward with the friends array [n_channels][n_prt] (incoming bincodes, 0 = none,
// whizard_gpu_phs.cpp:79-87); NULL = no friends.  A branch with a friend evaluates its
// cos-theta map at s = 4 E1 E2 (sqrts^2 when E1 / E2 are NULL).
int oracle_forward_friends(int n_in, int n_out, int n_channels, const int* d1, const int* d2, const int* hc,
                           const int* friends, const int* map_ids, const double* map_m, const double* map_w,
                           const double* flv, double sqrts, const double* E1, const double* E2, long n_events,
                           const int* channels, int n_x, const double* x, double* p, double* factors,
                           double* volumes, unsigned char* oks, double* msq, double* scale, double* cond) {
  Proc P;
  if (!build_proc(P, n_in, n_out, n_channels, d1, d2, hc, map_ids, flv)) return 1;
  if (friends) set_friends(P, friends);
  if (n_x < 3 * n_out - 4) return 2;
  MapPar<double> M;
  size_t nb = (size_t)n_channels * P.n_prt_out;
  M.m.assign(map_m, map_m + nb); M.w.assign(map_w, map_w + nb);
  M.a.assign(3 * nb, 0.0); M.b.assign(3 * nb, 0.0);
  double cur_sqrts = sqrts;
  for (int c = 0; c < n_channels; c++) init_constants(P, M, c, cur_sqrts);
  Ev<double> E;
  for (long i = 0; i < n_events; i++) {
    int c = channels[i];
    if (c < 0 || c >= n_channels) return 3;
    if (g_pert_seed) {   // this event's perturbation stream (per-event beams: their constants too)
      uint64_t st = g_pert_seed ^ (0xD1B54A32D192ED03ull * (uint64_t)(i + 1));
      g_pert = splitmix64(st) | 1;
    }
    double s_ev = sqrts, bg_lab = 0, s_friend = sqrts * sqrts;
    if (E1 && E2) {
      s_friend = 4 * E1[i] * E2[i];
      // phs.cu:1313-1317 (per-event sqrt(s)) and 1289 (lab boost parameter).
      s_ev = 2 * std::sqrt(E1[i] * E2[i]);
      bg_lab = (E1[i] - E2[i]) / (2 * std::sqrt(E1[i] * E2[i]));
      if (s_ev != cur_sqrts) {
        cur_sqrts = s_ev;
        for (int cc = 0; cc < n_channels; cc++) init_constants(P, M, cc, cur_sqrts);
      }
    }
    double f = 0, v = 0;
    std::memset(&p[4 * n_out * i], 0, 4 * n_out * sizeof(double));
    if (scale) std::memset(&scale[4 * n_out * i], 0, 4 * n_out * sizeof(double));
    bool ok = one_event<double>(P, M, c, s_ev, bg_lab, s_friend, &x[(size_t)n_x * i], E, &f, &v, &p[4 * n_out * i],
                                msq ? &msq[(size_t)P.n_prt_out * i] : nullptr,
                                scale ? &scale[4 * n_out * i] : nullptr, cond ? &cond[i] : nullptr);
    g_pert = 0;
    factors[i] = f; volumes[i] = v; oks[i] = ok ? 1 : 0;
  }
  return 0;
}

// Perturbation streams of oracle_forward (see pert()): 0 turns them off.
void oracle_set_perturbation(uint64_t seed) { g_pert_seed = seed; }

// Forward-mode derivatives (dual numbers) of the same generation, K <= 24 tangent
// directions: x_tan [n_events][n_x][K], m_tan / w_tan [n_channels][n_prt_out][K] seed the
// directions of the randoms and of the map masses / widths.  Outputs p_tan
// [n_events][n_out][4][K] and factor_tan [n_events][K] (zero for non-ok events), plus the
// primal values as oracle_forward.  friends (as oracle_forward_friends) and per-event beam
// energies E1, E2 may be NULL; with them the map constants are formed per event from the dual
// masses and widths, so their tangents follow each event's sqrt(s).
int oracle_forward_dual(int n_in, int n_out, int n_channels, const int* d1, const int* d2, const int* hc,
                        const int* friends, const int* map_ids, const double* map_m, const double* map_w,
                        const double* flv, double sqrts, const double* E1, const double* E2, long n_events,
                        const int* channels, int n_x, const double* x, int K, const double* x_tan,
                        const double* m_tan, const double* w_tan, double* p, double* factors, unsigned char* oks,
                        double* p_tan, double* factor_tan) {
  if (K < 1 || K > KD) return 4;
  Proc P;
  if (!build_proc(P, n_in, n_out, n_channels, d1, d2, hc, map_ids, flv)) return 1;
  if (friends) set_friends(P, friends);
  MapPar<Dual> M;
  size_t nb = (size_t)n_channels * P.n_prt_out;
  M.m.resize(nb); M.w.resize(nb); M.a.resize(3 * nb); M.b.resize(3 * nb);
  for (size_t i = 0; i < nb; i++) {
    M.m[i] = Dual(map_m[i]); M.w[i] = Dual(map_w[i]);
    for (int k = 0; k < K; k++) {
      M.m[i].d[k] = m_tan ? m_tan[i * K + k] : 0.0;
      M.w[i].d[k] = w_tan ? w_tan[i * K + k] : 0.0;
    }
  }
  double cur_sqrts = sqrts;
  for (int c = 0; c < n_channels; c++) init_constants(P, M, c, cur_sqrts);
  Ev<Dual> E;
  std::vector<Dual> xd(n_x), pd(4 * n_out);
  for (long i = 0; i < n_events; i++) {
    int c = channels[i];
    if (c < 0 || c >= n_channels) return 3;
    double s_ev = sqrts, bg_lab = 0, s_friend = sqrts * sqrts;
    if (E1 && E2) {   // as oracle_forward_friends
      s_friend = 4 * E1[i] * E2[i];
      s_ev = 2 * std::sqrt(E1[i] * E2[i]);
      bg_lab = (E1[i] - E2[i]) / (2 * std::sqrt(E1[i] * E2[i]));
      if (s_ev != cur_sqrts) {
        cur_sqrts = s_ev;
        for (int cc = 0; cc < n_channels; cc++) init_constants(P, M, cc, cur_sqrts);
      }
    }
    for (int k = 0; k < n_x; k++) {
      xd[k] = Dual(x[(size_t)n_x * i + k]);
      for (int t = 0; t < K; t++) xd[k].d[t] = x_tan ? x_tan[((size_t)n_x * i + k) * K + t] : 0.0;
    }
    for (auto& q : pd) q = Dual(0.0);
    Dual f, v;
    bool ok = one_event<Dual>(P, M, c, s_ev, bg_lab, s_friend, xd.data(), E, &f, &v, pd.data(), nullptr, nullptr, nullptr);
    oks[i] = ok ? 1 : 0;
    factors[i] = ok ? f.v : 0;
    for (int j = 0; j < 4 * n_out; j++) {
      p[4 * n_out * i + j] = ok ? pd[j].v : 0;
      for (int t = 0; t < K; t++) p_tan[((size_t)4 * n_out * i + j) * K + t] = ok ? pd[j].d[t] : 0;
    }
    for (int t = 0; t < K; t++) factor_tan[(size_t)K * i + t] = ok ? f.d[t] : 0;
  }
  return 0;
}

// Inverse mapping of n_events events into every channel (phs.cu:1499-1554): the randoms
// x_out [n_channels][n_events][n_x] (channel-major, as the Whizard interface returns them,
// phs.cu:222, 1585) and the factors [n_events][n_channels] each channel would give the event.
// p [n_events][n_out][4] are the momenta as the forward pass returns them (lab frame when
// E1 != E2; they are boosted back first).  Every channel is inverted, the generating one too.
// friends may be NULL.  Returns 0 on success.
int oracle_inverse(int n_in, int n_out, int n_channels, const int* d1, const int* d2, const int* hc,
                   const int* friends, const int* map_ids, const double* map_m, const double* map_w,
                   const double* flv, double sqrts, const double* E1, const double* E2, long n_events,
                   int n_x, const double* p, double* factors, double* x_out) {
  Proc P;
  if (!build_proc(P, n_in, n_out, n_channels, d1, d2, hc, map_ids, flv)) return 1;
  if (n_x < 3 * n_out - 4) return 2;
  if (friends) set_friends(P, friends);
  MapPar<double> M;
  size_t nb = (size_t)n_channels * P.n_prt_out;
  M.m.assign(map_m, map_m + nb); M.w.assign(map_w, map_w + nb);
  M.a.assign(3 * nb, 0.0); M.b.assign(3 * nb, 0.0);
  double cur_sqrts = sqrts;
  for (int c = 0; c < n_channels; c++) init_constants(P, M, c, cur_sqrts);
  InvState S;
  S.P.assign(4 * (size_t)P.n_prt, 0.0); S.msq.assign(P.n_prt, 0.0); S.pd.assign(P.n_prt, 0.0);
  std::vector<double> xb(n_x);
  for (long i = 0; i < n_events; i++) {
    double s_ev = sqrts, bg_lab = 0, s_friend = sqrts * sqrts;
    if (E1 && E2) {
      s_ev = 2 * std::sqrt(E1[i] * E2[i]);
      bg_lab = (E1[i] - E2[i]) / (2 * std::sqrt(E1[i] * E2[i]));
      s_friend = 4 * E1[i] * E2[i];
      if (s_ev != cur_sqrts) {
        cur_sqrts = s_ev;
        for (int cc = 0; cc < n_channels; cc++) init_constants(P, M, cc, cur_sqrts);
      }
    }
    const double g_lab = std::sqrt(1 + bg_lab * bg_lab);
    for (int c = 0; c < n_channels; c++) {
      // leaves: the event's momenta, back in the centre-of-mass frame (inverse of phs.cu:1286-1301)
      for (int j = 0, b = 1; j < n_out; j++, b *= 2) {
        const double* q = &p[(4 * (size_t)n_out) * i + 4 * j];
        double* Pl = &S.P[4 * (b - 1)];
        Pl[0] = q[0]; Pl[1] = q[1]; Pl[2] = q[2]; Pl[3] = q[3];
        if (bg_lab != 0) {
          Pl[0] = g_lab * q[0] - bg_lab * q[3];
          Pl[3] = -bg_lab * q[0] + g_lab * q[3];
        }
      }
      combine(P, c, P.root, S);
      for (int b = 0; b < P.n_prt_out; b++) {
        const double* q = &S.P[4 * b];
        S.msq[b] = q[0] * q[0] - q[1] * q[1] - q[2] * q[2] - q[3] * q[3];
      }
      S.pd[P.root] = 0;
      double f = 1;
      int id_x = 0;
      double m_tot = P.mass_sum[c * P.n_prt_out + P.root];
      inv_msq(P, M, c, P.root, s_ev, m_tot, S, xb.data(), &id_x, &f);
      double G[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
      inv_angles(P, M, c, P.root, s_friend, S, G, xb.data(), &id_x, &f);
      factors[(size_t)n_channels * i + c] = f;
      for (int k = 0; k < n_x; k++) x_out[((size_t)c * n_events + i) * n_x + k] = k < id_x ? xb[k] : 0.0;
    }
  }
  return 0;
}

}  // extern "C"
