// ORACLE — test infrastructure only.
//
// Driver around the reference's own CPU algorithm, compiled from the sources where they
// lie (phs_cpu.cpp, global_phs.cpp under the reference tree) by oracle/Makefile into
// oracle/_ref/libphs_ref.so.  Nothing here restates the algorithm: the harness fills the
// reference's process-global inputs exactly as the reference's own drivers do and calls
// the reference functions.  It is used (a) to write and check golden vectors for the
// oracle restatement and (b) as the timed CPU baseline ("kind": "reference").
//
// What the harness defines itself, and why:
//  * daughters1, daughters2, has_children, friends: phs.h:39-42 declares these tree tables
//    extern as phs_cpu.cpp's input; the reference defines them in phs.cu:27-30, a CUDA
//    file, and fills them in file_input.cpp:143-210 / whizard_gpu_phs.cpp:79-87.  They are
//    plain data that every driver fills.
//  * mass sums: phs.cu:61-78 (set_mass_sum) lives in the CUDA file too; the harness fills
//    mappings_host[c].mass_sum with the same recursion.
//  * The per-event loop body of phs_cpu.cpp:582-600 and 629-635 (resets, init_msq_cpu,
//    set_msq_cpu, set_angles_cpu), i.e. gen_phs_from_x_cpu_time_and_check with check_cpu and
//    do_inverse_mapping off.  That function itself is not linked: it reads input_control,
//    which the reference defines in file_input.cpp next to its JSON reader (rapidjson is
//    not in the image), so the link drops it (--gc-sections) instead of standing in for it.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "phs.h"
#include "global_phs.h"

int** daughters1 = nullptr;
int** daughters2 = nullptr;
int** has_children = nullptr;
int** friends = nullptr;

// Reference functions phs.h does not declare (phs_cpu.cpp:174, 283, 541).
void set_msq_cpu(int channel, int branch_idx, double m_tot, double* x, size_t* id_x, double sqrts,
                 phs_prt_t* prt, double* factor, double* volume, bool* ok, double* p_decay);
void set_angles_cpu(int i_event, int channel, int branch_idx, double* x, size_t* idx_x, double s,
                    double* factor, double* p_decay, phs_prt_t* prt, double L0[4][4]);
void init_msq_cpu(phs_prt_t* prt);

namespace {

void mass_sum_rec(int c, double* ms, int b) {  // phs.cu:61-78
  if (has_children[c][b]) {
    int k1 = daughters1[c][b], k2 = daughters2[c][b];
    mass_sum_rec(c, ms, k1);
    mass_sum_rec(c, ms, k2);
    ms[b] = ms[k1] + ms[k2];
  } else {
    int ld2 = 0, bb = b + 1;
    while (bb > 1) { bb /= 2; ld2++; }
    ms[b] = flv_masses[ld2 + 2];
  }
}

void release() {
  if (mappings_host) {
    for (int c = 0; c < N_CHANNELS; c++) {
      free(mappings_host[c].map_id); free(mappings_host[c].a); free(mappings_host[c].b);
      free(mappings_host[c].masses); free(mappings_host[c].widths); free(mappings_host[c].mass_sum);
      free(daughters1[c]); free(daughters2[c]); free(has_children[c]); free(friends[c]);
    }
    free(mappings_host); free(daughters1); free(daughters2); free(has_children); free(friends);
    free(flv_masses); free(flv_widths);
    mappings_host = nullptr;
  }
}

// Sizes as main.cpp:271-285; tables as main.cpp:295-320 and init_phs_gpu's index fix-up
// (phs.cu:392-397: daughters become bincode - 1).
int setup(int n_in, int n_out, int n_channels, const int* d1, const int* d2, const int* hc,
          const int* map_ids, const double* map_m, const double* map_w, const double* flv, double sqrts) {
  release();
  if (n_in != 2 || n_out < 2 || n_out > 12) return 1;
  N_CHANNELS = n_channels;
  N_EXT_IN = n_in; N_EXT_OUT = n_out; N_EXT_TOT = n_in + n_out;
  N_PRT = (1 << N_EXT_TOT) - 1;
  N_PRT_OUT = (1 << N_EXT_OUT) - 1;
  PRT_STRIDE = 4 * N_PRT;
  ROOT_BRANCH = N_PRT_OUT - 1;
  N_BRANCHES = 2 * N_EXT_OUT - 1;
  N_BRANCHES_INTERNAL = N_BRANCHES - N_EXT_OUT;
  N_MSQ = N_EXT_OUT - 2;
  N_LAMBDA_IN = N_EXT_OUT - 1;
  N_BOOSTS = N_LAMBDA_IN + 1;
  N_LAMBDA_OUT = N_BRANCHES;
  flv_masses = (double*)malloc(N_EXT_TOT * sizeof(double));
  flv_widths = (double*)calloc(N_EXT_TOT, sizeof(double));
  for (int i = 0; i < N_EXT_TOT; i++) flv_masses[i] = flv[i];
  daughters1 = (int**)malloc(n_channels * sizeof(int*));
  daughters2 = (int**)malloc(n_channels * sizeof(int*));
  has_children = (int**)malloc(n_channels * sizeof(int*));
  friends = (int**)malloc(n_channels * sizeof(int*));
  mappings_host = (mapping_t*)malloc(n_channels * sizeof(mapping_t));
  for (int c = 0; c < n_channels; c++) {
    daughters1[c] = (int*)malloc(N_PRT * sizeof(int));
    daughters2[c] = (int*)malloc(N_PRT * sizeof(int));
    has_children[c] = (int*)malloc(N_PRT * sizeof(int));
    friends[c] = (int*)calloc(N_PRT, sizeof(int));
    for (int i = 0; i < N_PRT; i++) {
      daughters1[c][i] = d1[c * N_PRT + i] - 1;
      daughters2[c][i] = d2[c * N_PRT + i] - 1;
      has_children[c][i] = hc[c * N_PRT + i];
    }
    mapping_t& m = mappings_host[c];
    m.map_id = (int*)malloc(N_PRT_OUT * sizeof(int));
    m.masses = (double*)malloc(N_PRT_OUT * sizeof(double));
    m.widths = (double*)malloc(N_PRT_OUT * sizeof(double));
    m.a = (map_constant_t*)malloc(N_PRT_OUT * sizeof(map_constant_t));
    m.b = (map_constant_t*)malloc(N_PRT_OUT * sizeof(map_constant_t));
    m.mass_sum = (double*)calloc(N_PRT_OUT, sizeof(double));
    m.comp_msq = nullptr; m.comp_ct = nullptr; m.comp_msq_inv = nullptr; m.comp_ct_inv = nullptr;
    for (int i = 0; i < N_PRT_OUT; i++) {
      m.map_id[i] = map_ids[c * N_PRT_OUT + i];
      m.masses[i] = map_m[c * N_PRT_OUT + i];
      m.widths[i] = map_w[c * N_PRT_OUT + i];
    }
    mass_sum_rec(c, m.mass_sum, ROOT_BRANCH);
  }
  init_mapping_constants_cpu(n_channels, sqrts);
  return 0;
}

struct Scratch {
  double* p_decay; double* msq; phs_prt_t* prt;
  Scratch() {
    p_decay = (double*)malloc(N_PRT * sizeof(double));
    msq = (double*)malloc(N_PRT * sizeof(double));
    prt = (phs_prt_t*)malloc(N_PRT * sizeof(phs_prt_t));
    memset(prt, 0, N_PRT * sizeof(phs_prt_t));
  }
  ~Scratch() { free(p_decay); free(msq); free(prt); }
};

// One event: phs_cpu.cpp:583-600 and 629-635.
inline bool event(Scratch& S, double sqrts, int c, double* x, double* factor, double* volume, double* msq_out) {
  bool ok = true;
  for (int j = 0; j < N_PRT; j++) S.msq[j] = -1;
  memset(S.p_decay, 0, N_PRT * sizeof(double));
  memset(S.prt, 0, 5 * N_PRT * sizeof(double));
  init_msq_cpu(S.prt);
  size_t id_x = 0;
  double m_tot = mappings_host[c].mass_sum[ROOT_BRANCH];
  double L0[4][4];
  memset(L0, 0, sizeof L0);
  L0[0][0] = L0[1][1] = L0[2][2] = L0[3][3] = 1;
  set_msq_cpu(c, ROOT_BRANCH, m_tot, x, &id_x, sqrts, S.prt, factor, volume, &ok, S.p_decay);
  if (msq_out)
    for (int i = 0; i < N_PRT_OUT; i++) msq_out[i] = S.prt[i].p2;
  if (ok) set_angles_cpu(0, c, ROOT_BRANCH, x, &id_x, sqrts * sqrts, factor, S.p_decay, S.prt, L0);
  for (int j = 0; j < N_PRT; j++)
    if (S.msq[j] < 0)
      S.msq[j] = S.prt[j].p[0] * S.prt[j].p[0] - S.prt[j].p[1] * S.prt[j].p[1] -
                 S.prt[j].p[2] * S.prt[j].p[2] - S.prt[j].p[3] * S.prt[j].p[3];
  return ok;
}

}  // namespace

extern "C" {

// Same argument convention as oracle_forward (fixed sqrt(s), no lab boost).
int ref_forward(int n_in, int n_out, int n_channels, const int* d1, const int* d2, const int* hc,
                const int* map_ids, const double* map_m, const double* map_w, const double* flv,
                double sqrts, long n_events, const int* channels, int n_x, double* x,
                double* p, double* factors, double* volumes, unsigned char* oks, double* msq) {
  int rc = setup(n_in, n_out, n_channels, d1, d2, hc, map_ids, map_m, map_w, flv, sqrts);
  if (rc) return rc;
  Scratch S;
  for (long i = 0; i < n_events; i++) {
    double f = 0, v = 0;
    bool ok = event(S, sqrts, channels[i], x + (size_t)n_x * i, &f, &v, msq ? msq + (size_t)N_PRT_OUT * i : nullptr);
    factors[i] = f; volumes[i] = v; oks[i] = ok;
    for (int j = 0, b = 1; j < n_out; j++, b *= 2)
      for (int mu = 0; mu < 4; mu++) p[(size_t)4 * n_out * i + 4 * j + mu] = ok ? S.prt[b - 1].p[mu] : 0.0;
  }
  return 0;
}

// Wall time (steady_clock, seconds) of the reference's generation loop over n_events
// (phs_cpu.cpp:582-667 with check_cpu and do_inverse_mapping off), single thread.
// Returns the number of ok events through *n_ok.
double ref_time_forward(int n_in, int n_out, int n_channels, const int* d1, const int* d2, const int* hc,
                        const int* map_ids, const double* map_m, const double* map_w, const double* flv,
                        double sqrts, long n_events, const int* channels, int n_x, double* x, long* n_ok) {
  if (setup(n_in, n_out, n_channels, d1, d2, hc, map_ids, map_m, map_w, flv, sqrts)) return -1.0;
  Scratch S;
  long ok_count = 0;
  double f, v;
  auto t0 = std::chrono::steady_clock::now();
  for (long i = 0; i < n_events; i++)
    ok_count += event(S, sqrts, channels[i], x + (size_t)n_x * i, &f, &v, nullptr);
  auto t1 = std::chrono::steady_clock::now();
  *n_ok = ok_count;
  return std::chrono::duration<double>(t1 - t0).count();
}

// The same loop on `threads` host threads, each over a contiguous range of the events with its
// own scratch (the reference's per-event functions write only to the arrays they are given and
// read the mapping tables, which setup() fills before the threads start). Wall time in seconds.
double ref_time_forward_mt(int n_in, int n_out, int n_channels, const int* d1, const int* d2, const int* hc,
                           const int* map_ids, const double* map_m, const double* map_w, const double* flv,
                           double sqrts, long n_events, const int* channels, int n_x, double* x, int threads,
                           long* n_ok) {
  if (threads < 1) return -1.0;
  if (setup(n_in, n_out, n_channels, d1, d2, hc, map_ids, map_m, map_w, flv, sqrts)) return -1.0;
  std::vector<long> oks(threads, 0);
  std::vector<std::thread> pool;
  auto t0 = std::chrono::steady_clock::now();
  for (int t = 0; t < threads; t++)
    pool.emplace_back([&, t] {
      Scratch S;
      double f, v;
      const long lo = n_events * t / threads, hi = n_events * (t + 1) / threads;
      long c = 0;
      for (long i = lo; i < hi; i++) c += event(S, sqrts, channels[i], x + (size_t)n_x * i, &f, &v, nullptr);
      oks[t] = c;
    });
  for (auto& th : pool) th.join();
  auto t1 = std::chrono::steady_clock::now();
  long ok_count = 0;
  for (long c : oks) ok_count += c;
  *n_ok = ok_count;
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
