// C ABI of the adaptive grids (include/gpu_phs.h): the reference's rng.cpp:18-93 (init_rng,
// rng_generate, update_weights) for device-resident events.
#include <cstring>
#include <map>
#include <memory>

#include "phs_state.h"

using namespace gphs;

namespace {

struct GridObj {
  int process = -1, n_channels = 0, n_dim = 0;
  std::vector<double> w;           // [channel][dim][kGridBins] bin weights
  double* cdf_d = nullptr;         // [plan channel][dim][kGridBins + 1]
  unsigned long long* counts_d = nullptr;
  size_t cdf_cap = 0, counts_cap = 0;
  ~GridObj() {
    if (cdf_d) (void)hipFree(cdf_d);
    if (counts_d) (void)hipFree(counts_d);
  }
};

std::map<int, std::unique_ptr<GridObj>>& grids() {
  static std::map<int, std::unique_ptr<GridObj>> g;
  return g;
}
int next_grid = 1;

GridObj* find_grid(int h) {
  auto it = grids().find(h);
  return it == grids().end() ? nullptr : it->second.get();
}

// The plan's segments must all come from the grid's process.
int check_plan(const GridObj& g, const PlanObj& pl) {
  for (int p : pl.chan_process)
    if (p != g.process) return set_error(GPU_PHS_E_ARG, "grid: the plan's segments belong to another process");
  return GPU_PHS_OK;
}

}  // namespace

extern "C" {

int gpu_phs_grid_create(int process) {
  std::lock_guard<std::mutex> lk(api_mutex());
  ProcessObj* po = find_process(process);
  if (!po) return set_error(GPU_PHS_E_HANDLE, "unknown process");
  auto g = std::make_unique<GridObj>();
  g->process = process;
  g->n_channels = (int)po->cp.channels.size();
  g->n_dim = po->cp.spec.n_x();
  // rng.cpp:20-27: equal bins of equal weight
  g->w.assign((size_t)g->n_channels * g->n_dim * kGridBins, 1.0 / kGridBins);
  const int h = next_grid++;
  grids()[h] = std::move(g);
  ok_status();
  return h;
}

int gpu_phs_grid_destroy(int grid) {
  std::lock_guard<std::mutex> lk(api_mutex());
  if (!grids().erase(grid)) return set_error(GPU_PHS_E_HANDLE, "unknown grid");
  return ok_status();
}

int gpu_phs_grid_get_weights(int grid, double* w, int n) {
  std::lock_guard<std::mutex> lk(api_mutex());
  GridObj* g = find_grid(grid);
  if (!g) return set_error(GPU_PHS_E_HANDLE, "unknown grid");
  if (!w || n < (int)g->w.size()) return set_error(GPU_PHS_E_ARG, "grid weights: array too short");
  std::memcpy(w, g->w.data(), g->w.size() * sizeof(double));
  return ok_status();
}

int gpu_phs_grid_set_weights(int grid, const double* w, int n) {
  std::lock_guard<std::mutex> lk(api_mutex());
  GridObj* g = find_grid(grid);
  if (!g) return set_error(GPU_PHS_E_HANDLE, "unknown grid");
  if (!w || n != (int)g->w.size()) return set_error(GPU_PHS_E_ARG, "grid weights: wrong length");
  for (int c = 0; c < g->n_channels * g->n_dim; c++) {
    double sum = 0;
    for (int b = 0; b < kGridBins; b++) {
      const double v = w[c * kGridBins + b];
      if (!(v >= 0)) return set_error(GPU_PHS_E_ARG, "grid weights must be >= 0");
      sum += v;
    }
    if (!(sum > 0)) return set_error(GPU_PHS_E_ARG, "grid weights: every dimension needs a positive weight");
  }
  g->w.assign(w, w + n);
  return ok_status();
}

int gpu_phs_grid_generate(int grid, int plan, uint64_t seed, void* stream) {
  std::lock_guard<std::mutex> lk(api_mutex());
  GridObj* g = find_grid(grid);
  if (!g) return set_error(GPU_PHS_E_HANDLE, "unknown grid");
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  int rc = check_plan(*g, *pl);
  if (rc) return rc;
  const int nb1 = kGridBins + 1;
  std::vector<double> cdf(pl->chans.size() * g->n_dim * nb1);
  for (size_t ci = 0; ci < pl->chans.size(); ci++) {
    const int pc = pl->chan_channel[ci];
    for (int d = 0; d < g->n_dim; d++) {
      const double* w = &g->w[((size_t)pc * g->n_dim + d) * kGridBins];
      double* F = &cdf[(ci * g->n_dim + d) * nb1];
      double sum = 0;
      for (int b = 0; b < kGridBins; b++) sum += w[b];
      F[0] = 0;
      for (int b = 0; b < kGridBins; b++) F[b + 1] = F[b] + w[b] / sum;
      F[kGridBins] = 1;
    }
  }
  hipError_t e;
  // the table of the previous call may still be read on this stream
  if ((e = hipStreamSynchronize((hipStream_t)stream)) != hipSuccess) return hip_error(e, "hipStreamSynchronize");
  if (cdf.size() > g->cdf_cap) {
    if (g->cdf_d) (void)hipFree(g->cdf_d);
    g->cdf_d = nullptr;
    g->cdf_cap = 0;
    if ((e = hipMalloc(&g->cdf_d, cdf.size() * sizeof(double))) != hipSuccess) return hip_error(e, "hipMalloc(grid)");
    g->cdf_cap = cdf.size();
  }
  if ((e = hipMemcpy(g->cdf_d, cdf.data(), cdf.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_error(e, "hipMemcpy(grid)");
  if ((e = launch_grid_generate(pl->dp, g->cdf_d, g->n_dim, seed, (hipStream_t)stream)) != hipSuccess)
    return hip_error(e, "grid generate launch");
  return ok_status();
}

int gpu_phs_grid_update(int grid, int plan, void* stream) {
  std::lock_guard<std::mutex> lk(api_mutex());
  GridObj* g = find_grid(grid);
  if (!g) return set_error(GPU_PHS_E_HANDLE, "unknown grid");
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  int rc = check_plan(*g, *pl);
  if (rc) return rc;
  const size_t nch = pl->chans.size(), nb = (size_t)g->n_dim * kGridBins;
  const size_t need = nch * nb + nch;
  hipError_t e;
  if (need > g->counts_cap) {
    if ((e = hipStreamSynchronize((hipStream_t)stream)) != hipSuccess) return hip_error(e, "hipStreamSynchronize");
    if (g->counts_d) (void)hipFree(g->counts_d);
    g->counts_d = nullptr;
    g->counts_cap = 0;
    if ((e = hipMalloc(&g->counts_d, need * sizeof(unsigned long long))) != hipSuccess)
      return hip_error(e, "hipMalloc(grid counts)");
    g->counts_cap = need;
  }
  if ((e = launch_grid_histogram(pl->dp, (int)nch, g->n_dim, g->counts_d, (hipStream_t)stream)) != hipSuccess)
    return hip_error(e, "grid histogram launch");
  std::vector<unsigned long long> h(need);
  if ((e = hipMemcpyAsync(h.data(), g->counts_d, need * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                          (hipStream_t)stream)) != hipSuccess ||
      (e = hipStreamSynchronize((hipStream_t)stream)) != hipSuccess)
    return hip_error(e, "grid counts download");
  // rng.cpp:48-64: weight = accepted events in the bin / accepted events of the channel; a
  // channel with no accepted event keeps its weights (the reference divides by zero)
  for (size_t ci = 0; ci < nch; ci++) {
    const unsigned long long n_ok = h[nch * nb + ci];
    if (n_ok == 0) continue;
    const int pc = pl->chan_channel[ci];
    for (size_t k = 0; k < nb; k++) g->w[(size_t)pc * nb + k] = (double)h[ci * nb + k] / (double)n_ok;
  }
  return ok_status();
}

}  // extern "C"
