// The Whizard-facing C ABI (include/gpu_phs.h group 1): the reference's 14 c_whizard_* entry
// points with the same signatures and index conventions (whizard_gpu_phs.cpp:12-249), on top
// of the process/plan machinery of phs_api.cpp.  The reference keeps its state in process
// globals (global_phs.cpp:6-27, phs.cu:27-41); this file keeps the equivalent in one static
// module object, so the calls keep their order-dependent semantics.
#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "phs_state.h"

namespace gphs {
namespace {

struct WhizardModule {
  int n_channels = 0, n_in = 0, n_out = 0, n_x = 0;
  bool sized = false, maps_ready = false, trees_ready = false, initialized = false;
  ProcessSpec spec;                 // trees and maps as filled
  std::vector<int> channel_ids;     // from c_whizard_init_channel_ids
  int block = 256;
  // compiled process for the current beams
  int process = -1;
  double E1 = -1, E2 = -1;
  // device buffers kept between calls (grow-only)
  double* x_d = nullptr; double* p_d = nullptr; double* f_d = nullptr; double* v_d = nullptr;
  double* e_d = nullptr;             // beam energies of the last call: E1 [n], then E2 [n] (phs.cu:1361-1363)
  size_t last_n = 0;                 // events of the last generate call
  uint8_t* ok_d = nullptr;
  size_t cap = 0;                   // events the buffers hold
  int cap_nx = 0, cap_nout = 0;
  // inverse mapping outputs: factors [n][n_channels], randoms [n_channels][n][n_x]
  double* fa_d = nullptr; double* xo_d = nullptr;
  size_t inv_cap = 0;               // doubles of xo_d
  double* last_p = nullptr;         // device momenta of the last generate call
  // the plan of the last call, reused (tables, inverse tables) while its segments repeat
  std::unique_ptr<PlanObj> plan;
  std::vector<gpu_phs_segment> plan_segs;
  bool plan_varying = false;
  hipStream_t stream = nullptr;     // the module's own stream (non-blocking, default priority)
  // pinned staging of the host copies: two slots, so the host-side copy of one chunk overlaps
  // the DMA of the other
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
};

constexpr size_t kPinChunk = (size_t)16 << 20;

WhizardModule& W() {
  static WhizardModule w;
  return w;
}

void free_buffers(WhizardModule& w) {
  w.plan.reset();   // its segments point into the buffers
  w.plan_segs.clear();
  (void)hipFree(w.x_d); (void)hipFree(w.p_d); (void)hipFree(w.f_d); (void)hipFree(w.v_d); (void)hipFree(w.ok_d);
  (void)hipFree(w.fa_d); (void)hipFree(w.xo_d); (void)hipFree(w.e_d);
  w.x_d = w.p_d = w.f_d = w.v_d = nullptr;
  w.fa_d = w.xo_d = w.e_d = nullptr;
  w.last_n = 0;
  w.ok_d = nullptr;
  w.cap = 0;
  w.inv_cap = 0;
  w.last_p = nullptr;
}

int ensure_buffers(WhizardModule& w, size_t n) {
  if (n <= w.cap && w.cap_nx == w.n_x && w.cap_nout == w.n_out) return GPU_PHS_OK;
  free_buffers(w);
  hipError_t e;
  if ((e = hipMalloc(&w.x_d, n * w.n_x * sizeof(double))) != hipSuccess ||
      (e = hipMalloc(&w.p_d, n * w.n_out * 4 * sizeof(double))) != hipSuccess ||
      (e = hipMalloc(&w.f_d, n * sizeof(double))) != hipSuccess ||
      (e = hipMalloc(&w.v_d, n * sizeof(double))) != hipSuccess ||
      (e = hipMalloc(&w.e_d, 2 * n * sizeof(double))) != hipSuccess ||
      (e = hipMalloc(&w.ok_d, n)) != hipSuccess) {
    free_buffers(w);
    return hip_error(e, "hipMalloc(event buffers)");
  }
  w.cap = n; w.cap_nx = w.n_x; w.cap_nout = w.n_out;
  return GPU_PHS_OK;
}

int ensure_inverse_buffers(WhizardModule& w, size_t n) {
  const size_t need = n * w.n_channels * w.n_x;
  if (need <= w.inv_cap && w.fa_d) return GPU_PHS_OK;
  (void)hipFree(w.fa_d); (void)hipFree(w.xo_d);
  w.fa_d = w.xo_d = nullptr;
  w.inv_cap = 0;
  hipError_t e;
  if ((e = hipMalloc(&w.fa_d, n * w.n_channels * sizeof(double))) != hipSuccess ||
      (e = hipMalloc(&w.xo_d, std::max<size_t>(need, 1) * sizeof(double))) != hipSuccess) {
    (void)hipFree(w.fa_d); w.fa_d = nullptr;
    return hip_error(e, "hipMalloc(inverse buffers)");
  }
  w.inv_cap = need;
  return GPU_PHS_OK;
}

void free_staging(WhizardModule& w) {
  if (w.stream) (void)hipStreamSynchronize(w.stream);
  for (int k = 0; k < 2; k++) {
    if (w.pin[k]) (void)hipHostFree(w.pin[k]);
    if (w.pin_ev[k]) (void)hipEventDestroy(w.pin_ev[k]);
    w.pin[k] = nullptr;
    w.pin_ev[k] = nullptr;
  }
  if (w.stream) (void)hipStreamDestroy(w.stream);
  w.stream = nullptr;
}

int ensure_staging(WhizardModule& w) {
  if (w.stream) return GPU_PHS_OK;
  hipError_t e = hipStreamCreateWithFlags(&w.stream, hipStreamNonBlocking);
  if (e != hipSuccess) { w.stream = nullptr; return hip_error(e, "hipStreamCreate"); }
  for (int k = 0; k < 2; k++) {
    if ((e = hipHostMalloc(&w.pin[k], kPinChunk, hipHostMallocDefault)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&w.pin_ev[k], hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventRecord(w.pin_ev[k], w.stream)) != hipSuccess) {
      free_staging(w);
      return hip_error(e, "pinned staging buffers");
    }
  }
  return GPU_PHS_OK;
}

// Host copies of the staging chunks, each split over kCopyThreads threads: the calling thread
// and kCopyThreads - 1 workers started at the first large copy and kept until the library is
// unloaded.  One thread copies some 10-15 GB/s between pinned and pageable memory, well below
// what PCIe moves, so a single thread would set the entry point's rate; starting threads for
// every 16 MiB chunk would cost tens of microseconds per chunk.
constexpr int kCopyThreads = 8;
class CopyPool {
 public:
  void copy(void* dst, const void* src, size_t n) {
    if (n < ((size_t)1 << 20)) {
      std::memcpy(dst, src, n);
      return;
    }
    const size_t part = ((n + kCopyThreads - 1) / kCopyThreads + 63) & ~(size_t)63;
    {
      std::lock_guard<std::mutex> lk(m_);
      if (th_.empty())
        for (int k = 1; k < kCopyThreads; k++) th_.emplace_back([this, k] { worker(k); });
      dst_ = (char*)dst;
      src_ = (const char*)src;
      n_ = n;
      part_ = part;
      pending_ = kCopyThreads - 1;
      gen_++;
    }
    work_.notify_all();
    std::memcpy(dst, src, std::min(part, n));
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [&] { return pending_ == 0; });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    work_.notify_all();
    for (auto& t : th_) t.join();
  }

 private:
  void worker(int k) {
    uint64_t seen = 0;
    for (;;) {
      char* dst;
      const char* src;
      size_t n, part;
      {
        std::unique_lock<std::mutex> lk(m_);
        work_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        dst = dst_; src = src_; n = n_; part = part_;
      }
      const size_t off = part * k;
      if (off < n) std::memcpy(dst + off, src + off, std::min(part, n - off));
      std::lock_guard<std::mutex> lk(m_);
      if (--pending_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable work_, done_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t n_ = 0, part_ = 0;
  int pending_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

void chunk_copy(void* dst, const void* src, size_t n) {
  static CopyPool pool;
  pool.copy(dst, src, n);
}

// Host -> device through the pinned slots: each chunk is copied into a free slot on the host and
// sent by DMA on the module's stream while the next chunk is copied into the other slot.
int h2d(WhizardModule& w, void* dst, const void* src, size_t bytes) {
  int k = 0;
  for (size_t off = 0; off < bytes; off += kPinChunk, k ^= 1) {
    const size_t n = std::min(kPinChunk, bytes - off);
    hipError_t e = hipEventSynchronize(w.pin_ev[k]);   // the slot's previous DMA is done
    if (e != hipSuccess) return hip_error(e, "hipEventSynchronize");
    chunk_copy(w.pin[k], (const char*)src + off, n);
    if ((e = hipMemcpyAsync((char*)dst + off, w.pin[k], n, hipMemcpyHostToDevice, w.stream)) != hipSuccess ||
        (e = hipEventRecord(w.pin_ev[k], w.stream)) != hipSuccess)
      return hip_error(e, "hipMemcpyAsync(host to device)");
  }
  return GPU_PHS_OK;
}

// Device -> host (after the work queued on the module's stream): chunk i + 1 is sent by DMA into
// one slot while chunk i is copied out of the other on the host.
int d2h(WhizardModule& w, void* dst, const void* src, size_t bytes) {
  const size_t n_chunks = (bytes + kPinChunk - 1) / kPinChunk;
  auto issue = [&](size_t i) -> hipError_t {
    const size_t off = i * kPinChunk, n = std::min(kPinChunk, bytes - off);
    hipError_t e = hipMemcpyAsync(w.pin[i % 2], (const char*)src + off, n, hipMemcpyDeviceToHost, w.stream);
    return e != hipSuccess ? e : hipEventRecord(w.pin_ev[i % 2], w.stream);
  };
  hipError_t e = n_chunks ? issue(0) : hipSuccess;
  for (size_t i = 0; i < n_chunks && e == hipSuccess; i++) {
    if (i + 1 < n_chunks && (e = issue(i + 1)) != hipSuccess) break;
    if ((e = hipEventSynchronize(w.pin_ev[i % 2])) != hipSuccess) break;
    const size_t off = i * kPinChunk;
    chunk_copy((char*)dst + off, w.pin[i % 2], std::min(kPinChunk, bytes - off));
  }
  return e == hipSuccess ? GPU_PHS_OK : hip_error(e, "device to host copy");
}

int compile_for_beams(WhizardModule& w, double E1, double E2) {
  if (w.process >= 0 && E1 == w.E1 && E2 == w.E2) return GPU_PHS_OK;
  if (w.process >= 0) { drop_process(w.process); w.process = -1; }
  ProcessSpec sp = w.spec;
  sp.E1 = E1; sp.E2 = E2;
  auto po = std::make_unique<ProcessObj>();
  std::string err = compile_process(sp, po->cp);
  if (!err.empty()) return set_error(GPU_PHS_E_TREE, err);
  w.process = register_process(std::move(po));
  w.E1 = E1; w.E2 = E2;
  return GPU_PHS_OK;
}

}  // namespace
}  // namespace gphs

using namespace gphs;

extern "C" {

void c_whizard_set_particle_structure(int* n_channels, int* n_trees, int* n_groves, int* n_x, int* n_in, int* n_out) {
  std::lock_guard<std::mutex> lk(api_mutex());
  (void)n_trees; (void)n_groves;
  WhizardModule& w = W();
  if (!n_channels || !n_x || !n_in || !n_out || *n_channels < 1 || *n_in != 2 || *n_out < 2 || *n_out > kMaxOut) {
    set_error(GPU_PHS_E_ARG, "set_particle_structure: need n_channels >= 1, n_in = 2, 2 <= n_out <= 12");
    return;
  }
  w.n_channels = *n_channels; w.n_in = *n_in; w.n_out = *n_out; w.n_x = *n_x;
  w.spec = ProcessSpec();
  w.spec.n_in = w.n_in; w.spec.n_out = w.n_out; w.spec.n_channels = w.n_channels;
  w.spec.flv_masses.assign(w.n_in + w.n_out, 0.0);
  w.spec.flv_widths.assign(w.n_in + w.n_out, 0.0);
  w.sized = true; w.maps_ready = w.trees_ready = w.initialized = false;
  if (w.n_x != w.spec.n_x()) {
    set_error(GPU_PHS_E_ARG, "set_particle_structure: n_x must be 3 n_out - 4");
    return;
  }
  ok_status();
}

void c_whizard_set_flavors(double* masses, double* widths) {
  std::lock_guard<std::mutex> lk(api_mutex());
  WhizardModule& w = W();
  if (!w.sized) { set_error(GPU_PHS_E_STATE, "set_flavors before set_particle_structure"); return; }
  for (int i = 0; i < w.n_in + w.n_out; i++) {
    w.spec.flv_masses[i] = masses[i];
    w.spec.flv_widths[i] = widths ? widths[i] : 0.0;
  }
  if (w.process >= 0) { drop_process(w.process); w.process = -1; }
  ok_status();
}

void c_whizard_init_mappings(void) {
  std::lock_guard<std::mutex> lk(api_mutex());
  WhizardModule& w = W();
  if (!w.sized) { set_error(GPU_PHS_E_STATE, "init_mappings before set_particle_structure"); return; }
  size_t n = (size_t)w.n_channels * w.spec.n_prt_out();
  w.spec.map_ids.assign(n, 0);
  w.spec.map_masses.assign(n, 0.0);
  w.spec.map_widths.assign(n, 0.0);
  w.maps_ready = true;
  ok_status();
}

void c_whizard_fill_mapping(int* channel, int* map_ids, double* masses, double* widths) {
  std::lock_guard<std::mutex> lk(api_mutex());
  WhizardModule& w = W();
  if (!w.maps_ready) { set_error(GPU_PHS_E_STATE, "fill_mapping before init_mappings"); return; }
  int c = *channel - 1;  // Fortran -> C index
  if (c < 0 || c >= w.n_channels) { set_error(GPU_PHS_E_ARG, "fill_mapping: channel out of range"); return; }
  const int n = w.spec.n_prt_out();
  for (int i = 0; i < n; i++) {
    w.spec.map_ids[c * n + i] = map_ids[i];
    w.spec.map_masses[c * n + i] = masses[i];
    w.spec.map_widths[c * n + i] = widths[i];
  }
  ok_status();
}

void c_whizard_init_tree_structures(void) {
  std::lock_guard<std::mutex> lk(api_mutex());
  WhizardModule& w = W();
  if (!w.sized) { set_error(GPU_PHS_E_STATE, "init_tree_structures before set_particle_structure"); return; }
  size_t n = (size_t)w.n_channels * w.spec.n_prt();
  w.spec.d1.assign(n, 0); w.spec.d2.assign(n, 0); w.spec.has_children.assign(n, 0); w.spec.friends.assign(n, 0);
  w.trees_ready = true;
  ok_status();
}

void c_whizard_fill_tree_structure(int* channel, int* daughters1, int* daughters2, int* has_children, int* friends) {
  std::lock_guard<std::mutex> lk(api_mutex());
  WhizardModule& w = W();
  if (!w.trees_ready) { set_error(GPU_PHS_E_STATE, "fill_tree_structure before init_tree_structures"); return; }
  int c = *channel - 1;
  if (c < 0 || c >= w.n_channels) { set_error(GPU_PHS_E_ARG, "fill_tree_structure: channel out of range"); return; }
  const int n = w.spec.n_prt();
  for (int i = 0; i < n; i++) {
    w.spec.d1[c * n + i] = daughters1[i];
    w.spec.d2[c * n + i] = daughters2[i];
    w.spec.has_children[c * n + i] = has_children[i];
    w.spec.friends[c * n + i] = friends ? friends[i] : 0;
  }
  ok_status();
}

void c_whizard_init_channel_ids(int* batch_size, int* n_channels, int* channel_limits) {
  std::lock_guard<std::mutex> lk(api_mutex());
  WhizardModule& w = W();
  if (!batch_size || !n_channels || !channel_limits || *batch_size < 0 || *n_channels < 1) {
    set_error(GPU_PHS_E_ARG, "init_channel_ids: invalid arguments");
    return;
  }
  // Event i belongs to the first channel c with i < channel_limits[c] (cumulative ends).
  // The reference skips only leading empty channels (whizard_gpu_phs.cpp:93-97); every empty
  // channel is skipped here.
  w.channel_ids.assign(*batch_size, -1);
  int c = 0;
  for (int i = 0; i < *batch_size; i++) {
    while (c < *n_channels && channel_limits[c] <= i) c++;
    if (c >= *n_channels) {
      set_error(GPU_PHS_E_ARG, "init_channel_ids: channel_limits end before batch_size");
      w.channel_ids.clear();
      return;
    }
    w.channel_ids[i] = c;
  }
  ok_status();
}

void c_whizard_set_threads(int* msq_threads, int* cb_threads, int* ab_threads) {
  std::lock_guard<std::mutex> lk(api_mutex());
  (void)cb_threads; (void)ab_threads;
  int t = msq_threads ? *msq_threads : 256;
  t = (t / 64) * 64;
  if (t < 64) t = 64;
  if (t > 256) t = 256;
  W().block = t;
  ok_status();
}

void c_whizard_init_gpu_phs(void) {
  std::lock_guard<std::mutex> lk(api_mutex());
  WhizardModule& w = W();
  if (!w.sized || !w.maps_ready || !w.trees_ready) {
    set_error(GPU_PHS_E_STATE, "init_gpu_phs before the structure, mappings and trees were set");
    return;
  }
  // validate the trees now (constants are recompiled for the beams of each generate call)
  ProcessSpec sp = w.spec;
  sp.E1 = sp.E2 = 0.5 * 1e3;
  CompiledProcess cp;
  std::string err = compile_process(sp, cp);
  if (!err.empty()) { set_error(GPU_PHS_E_TREE, err); return; }
  if (require_device()) return;
  w.initialized = true;
  ok_status();
}

void c_whizard_finalize_gpu_phs(void) {
  std::lock_guard<std::mutex> lk(api_mutex());
  WhizardModule& w = W();
  free_buffers(w);
  free_staging(w);
  if (w.process >= 0) drop_process(w.process);
  w = WhizardModule();
  ok_status();
}

void c_whizard_gen_phs_from_x_gpu(double* E1_in, double* E2_in, bool* requires_boost_to_lab, int* n_events,
                                  int* n_channels, int* n_x, double* x, double* factors, double* volumes, bool* oks,
                                  double* p, double* x_out) {
  std::lock_guard<std::mutex> lk(api_mutex());
  (void)requires_boost_to_lab;
  WhizardModule& w = W();
  if (!w.initialized) { set_error(GPU_PHS_E_STATE, "gen_phs_from_x_gpu before init_gpu_phs"); return; }
  const int64_t n = *n_events;
  if (n <= 0) { ok_status(); return; }
  if (*n_x != w.n_x || *n_channels != w.n_channels) {
    set_error(GPU_PHS_E_ARG, "gen_phs_from_x_gpu: n_x / n_channels differ from set_particle_structure");
    return;
  }
  if ((int64_t)w.channel_ids.size() < n) {
    set_error(GPU_PHS_E_STATE, "gen_phs_from_x_gpu: init_channel_ids covers fewer events");
    return;
  }
  // beams: fixed, or per event (phs.cu:1313-1317); the trees are compiled for the first
  // event's, and with per-event beams the kernels recompute what depends on them
  const double E1 = E1_in[0], E2 = E2_in[0];
  bool varying = false;
  for (int64_t i = 0; i < n; i++) {
    if (!(E1_in[i] > 0) || !(E2_in[i] > 0)) {
      set_error(GPU_PHS_E_ARG, "gen_phs_from_x_gpu: beam energies must be positive");
      return;
    }
    if (E1_in[i] != E1 || E2_in[i] != E2) varying = true;
  }
  if (compile_for_beams(w, E1, E2)) return;
  if (w.plan && !w.plan_segs.empty() && w.plan_segs[0].process != w.process) { w.plan.reset(); w.plan_segs.clear(); }
  if (ensure_buffers(w, (size_t)n)) return;
  if (ensure_staging(w)) return;
  if (h2d(w, w.x_d, x, (size_t)n * w.n_x * sizeof(double))) return;
  // the beam energies stay on the device, E1 then E2, as the reference keeps them
  // (get_in_momentum_device_pointer, phs.cu:1361-1363); per-event beams read them from there
  if (h2d(w, w.e_d, E1_in, (size_t)n * sizeof(double)) || h2d(w, w.e_d + n, E2_in, (size_t)n * sizeof(double))) return;

  // segments = runs of equal channel id
  std::vector<gpu_phs_segment> segs;
  for (int64_t i = 0; i < n;) {
    int64_t j = i;
    while (j < n && w.channel_ids[j] == w.channel_ids[i]) j++;
    gpu_phs_segment s{};
    s.process = w.process;
    s.channel = w.channel_ids[i];
    s.n_events = j - i;
    s.x = w.x_d + i * w.n_x;
    s.p = w.p_d + i * w.n_out * 4;
    s.factors = w.f_d + i;
    s.volumes = w.v_d + i;
    s.oks = w.ok_d + i;
    segs.push_back(s);
    i = j;
  }
  // the plan (and its inverse tables) is kept while the segments and the beam mode repeat
  const bool same = w.plan && w.plan_varying == varying && w.plan_segs.size() == segs.size() &&
                    std::memcmp(w.plan_segs.data(), segs.data(), segs.size() * sizeof(gpu_phs_segment)) == 0;
  if (!same) {
    if (!w.plan) w.plan = std::make_unique<PlanObj>();
    PlanObj& pl = *w.plan;
    pl.user = segs;
    pl.beam_e1.clear();
    pl.beam_e2.clear();
    if (varying) {
      for (size_t k = 0, r = 0; k < segs.size(); r += segs[k].n_events, k++) {
        pl.beam_e1.push_back(w.e_d + r);
        pl.beam_e2.push_back(w.e_d + n + r);
      }
    }
    // each tile class keeps its default shape with the block capped at set_threads' value,
    // and smaller where the LDS tables of the deepest trees would not fit
    const PlanObj defaults;
    for (int c = 0; c < kTileClasses; c++) pl.lcs[c].block = std::min(defaults.lcs[c].block, w.block);
    if (plan_build(pl)) { w.plan.reset(); return; }
    for (int c = 0; c < kTileClasses; c++)
      while (pl.lcs[c].block > 64 && forward_lds_bytes(pl.dps[c], pl.lcs[c]) > forward_lds_limit()) {
        pl.lcs[c].block -= 64;
        if (plan_build(pl)) { w.plan.reset(); return; }
      }
    w.plan_segs = segs;
    w.plan_varying = varying;
  }
  PlanObj& pl = *w.plan;
  if (plan_run(pl, w.stream)) return;
  // the inverse mapping of every event into every channel (the reference always asks for it,
  // whizard_gpu_phs.cpp:118): factors [n][n_channels] at [n_channels i + c] (phs.cu:207) and
  // x_out channel-major [n_x n c + n_x i + k] (phs.cu:222), own channel included
  if (ensure_inverse_buffers(w, (size_t)n)) return;
  std::vector<int64_t> rows(segs.size());
  for (size_t k = 0, r = 0; k < segs.size(); r += segs[k].n_events, k++) rows[k] = (int64_t)r;
  if (plan_inverse(pl, rows.data(), n, w.fa_d, w.xo_d, 0, w.stream)) return;
  const int nc = w.n_channels;
  if (d2h(w, p, w.p_d, (size_t)n * w.n_out * 4 * sizeof(double)) ||
      d2h(w, factors, w.fa_d, (size_t)n * nc * sizeof(double)) ||
      d2h(w, volumes, w.v_d, (size_t)n * sizeof(double)) || d2h(w, oks, w.ok_d, (size_t)n) ||
      (x_out && d2h(w, x_out, w.xo_d, (size_t)n * nc * w.n_x * sizeof(double))))
    return;
  w.last_p = w.p_d;
  w.last_n = (size_t)n;
  ok_status();
}

void c_whizard_get_momentum_device_pointer(double* p) {
  std::lock_guard<std::mutex> lk(api_mutex());
  if (p) {
    uint64_t addr = (uint64_t)(uintptr_t)W().last_p;
    std::memcpy(p, &addr, sizeof addr);
  }
  ok_status();
}

}  // extern "C"

// transfer_to_whizard.h:1-3 (C++ linkage, as the reference declares them): device pointers of
// the last generate call's momenta ([n_events][n_out][4], the Whizard layout: the reference
// returns its internal per-branch array and gathers on the host, phs.cu:1566-1575), beam
// energies (E1 [n_events], then E2 [n_events]) and ok flags ([n_events] bool).  NULL before
// the first call.
double* get_momentum_device_pointer() {
  std::lock_guard<std::mutex> lk(api_mutex());
  return W().last_p;
}
double* get_in_momentum_device_pointer() {
  std::lock_guard<std::mutex> lk(api_mutex());
  return W().last_n ? W().e_d : nullptr;
}
bool* get_ok_device_pointer() {
  std::lock_guard<std::mutex> lk(api_mutex());
  return W().last_n ? reinterpret_cast<bool*>(W().ok_d) : nullptr;
}

extern "C" {

void c_whizard_reset_generator(void) {
  std::lock_guard<std::mutex> lk(api_mutex());
  WhizardModule& w = W();
  free_buffers(w);
  free_staging(w);
  std::fill(w.spec.flv_masses.begin(), w.spec.flv_masses.end(), 0.0);
  std::fill(w.spec.flv_widths.begin(), w.spec.flv_widths.end(), 0.0);
  if (w.process >= 0) { drop_process(w.process); w.process = -1; }
  ok_status();
}

void c_whizard_show_module(void) {
  std::lock_guard<std::mutex> lk(api_mutex());
  WhizardModule& w = W();
  const ProcessSpec& sp = w.spec;
  printf("n_channels: %d\nn_ext_in: %d\nn_ext_out: %d\nn_ext_tot: %d\n", w.n_channels, w.n_in, w.n_out, w.n_in + w.n_out);
  if (!w.sized) { printf(" STRUCTURE NOT SET UP!\n"); ok_status(); return; }
  printf("n_prt: %d\nn_prt_out: %d\nroot_branch: %d\nn_branches: %d\nn_x: %d\n", sp.n_prt(), sp.n_prt_out(), sp.root(),
         2 * w.n_out - 1, w.n_x);
  printf("flv_masses:");
  for (double m : sp.flv_masses) printf(" %lf", m);
  printf("\nMappings:\n");
  if (!w.maps_ready) printf(" MAPPINGS NOT SET UP!\n");
  else
    for (int c = 0; c < w.n_channels; c++) {
      printf(" c: %d\n ids:", c);
      for (int i = 0; i < sp.n_prt_out(); i++) printf(" %d", sp.map_ids[c * sp.n_prt_out() + i]);
      printf("\n masses:");
      for (int i = 0; i < sp.n_prt_out(); i++) printf(" %lf", sp.map_masses[c * sp.n_prt_out() + i]);
      printf("\n widths:");
      for (int i = 0; i < sp.n_prt_out(); i++) printf(" %lf", sp.map_widths[c * sp.n_prt_out() + i]);
      printf("\n");
    }
  printf("Trees:\n");
  if (!w.trees_ready) printf(" TREES NOT SET UP!\n");
  else
    for (int c = 0; c < w.n_channels; c++) {
      printf(" channel %d:", c);
      for (int i = 0; i < sp.n_prt(); i++)
        if (sp.has_children[c * sp.n_prt() + i])
          printf(" %d->(%d,%d)", i + 1, sp.d1[c * sp.n_prt() + i], sp.d2[c * sp.n_prt() + i]);
      printf("\n");
    }
  if (w.initialized && w.process >= 0) {
    ProcessObj* po = find_process(w.process);
    if (po) {
      printf("Programs (visit order: branch, phi random, daughters, next):\n");
      for (size_t c = 0; c < po->cp.channels.size(); c++) {
        const ChannelDesc& cd = po->cp.channels[c];
        printf(" channel %zu:", c);
        for (int i = 0; i < cd.n_nodes; i++) {
          const NodeCmd& nd = po->cp.nodes[cd.node0 + i];
          printf(" [%d x%d desc=%d push=%d]", po->cp.node_branch[cd.node0 + i] + 1, nd.x_phi, nd.desc, nd.push);
        }
        printf("\n");
      }
    }
  }
  printf("END OF MODULE\n");
  fflush(stdout);
  ok_status();
}

}  // extern "C"
