// Process handles, batched plans, launches, hipGraph replay and the autotuner
// (the additions of include/gpu_phs.h; no counterpart in the reference).
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <functional>
#include <numeric>

#include "phs_shapes.h"
#include "phs_state.h"

namespace gphs {

namespace {
thread_local int t_err = 0;
thread_local std::string t_msg;
std::map<int, std::unique_ptr<ProcessObj>> g_procs;
std::map<int, std::unique_ptr<PlanObj>> g_plans;
int g_next_proc = 0, g_next_plan = 0;
}  // namespace

int set_error(int code, const std::string& msg) {
  t_err = code;
  t_msg = msg;
  return code;
}

int hip_error(hipError_t e, const char* what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return set_error(GPU_PHS_E_HIP, m);
}

std::mutex& api_mutex() {
  static std::mutex m;
  return m;
}

int require_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) return set_error(GPU_PHS_E_NODEVICE, "no GPU device is visible");
  return GPU_PHS_OK;
}

int register_process(std::unique_ptr<ProcessObj> p) {
  int h = g_next_proc++;
  g_procs[h] = std::move(p);
  return h;
}
ProcessObj* find_process(int h) {
  auto it = g_procs.find(h);
  return it == g_procs.end() ? nullptr : it->second.get();
}
int drop_process(int h) { return g_procs.erase(h) ? GPU_PHS_OK : GPU_PHS_E_HANDLE; }
int register_plan(std::unique_ptr<PlanObj> p) {
  int h = g_next_plan++;
  g_plans[h] = std::move(p);
  return h;
}
PlanObj* find_plan(int h) {
  auto it = g_plans.find(h);
  return it == g_plans.end() ? nullptr : it->second.get();
}
int drop_plan(int h) { return g_plans.erase(h) ? GPU_PHS_OK : GPU_PHS_E_HANDLE; }

PlanObj::~PlanObj() {
  if (adj_dev) (void)hipFree(adj_dev);
  if (inv_dev) (void)hipFree(inv_dev);
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (capture_stream) (void)hipStreamDestroy(capture_stream);
  if (side_stream) (void)hipStreamDestroy(side_stream);
  if (fork_event) (void)hipEventDestroy(fork_event);
  if (join_event) (void)hipEventDestroy(join_event);
  for (int k = 0; k < kMaxStreams - 1; k++) {
    if (lane_streams[k]) (void)hipStreamDestroy(lane_streams[k]);
    if (lane_join[k]) (void)hipEventDestroy(lane_join[k]);
  }
  if (dev) (void)hipFree(dev);
}

static void drop_graph(PlanObj& pl) {
  if (pl.graph_exec) { (void)hipGraphExecDestroy(pl.graph_exec); pl.graph_exec = nullptr; }
  if (pl.graph) { (void)hipGraphDestroy(pl.graph); pl.graph = nullptr; }
}

// Tables of a plan: channels of every process the segments use (deduplicated), their
// programs, segment descriptors, and the tile table (tiles never cross a segment, so every
// workgroup runs a single channel).
int plan_build(PlanObj& pl) {
  pl.chans.clear(); pl.nodes.clear(); pl.segs.clear(); pl.tiles.clear();
  pl.seg_tile_begin.clear(); pl.seg_tile_end.clear();
  pl.chan_process.clear(); pl.chan_channel.clear();
  if (pl.adj_dev) { (void)hipFree(pl.adj_dev); pl.adj_dev = nullptr; pl.adj_bytes = 0; }
  pl.adj_valid = false;
  pl.n_banks = 0;
  std::map<std::pair<int, int>, int> chan_index;
  for (const auto& us : pl.user) {
    ProcessObj* po = find_process(us.process);
    if (!po) return set_error(GPU_PHS_E_HANDLE, "segment refers to an unknown process");
    const CompiledProcess& cp = po->cp;
    if (us.channel < 0 || us.channel >= (int)cp.channels.size())
      return set_error(GPU_PHS_E_ARG, "segment channel out of range");
    if (us.n_events < 0) return set_error(GPU_PHS_E_ARG, "negative event count");
    if (us.n_events > 0 && (!us.x || !us.p || !us.factors || !us.volumes || !us.oks))
      return set_error(GPU_PHS_E_ARG, "segment buffers must not be NULL");
    auto key = std::make_pair(us.process, us.channel);
    int ci;
    auto it = chan_index.find(key);
    if (it == chan_index.end()) {
      ChannelDesc cd = cp.channels[us.channel];
      int node0 = (int)pl.nodes.size();
      for (int i = 0; i < cd.n_nodes; i++) pl.nodes.push_back(cp.nodes[cd.node0 + i]);
      cd.node0 = node0;
      cd.param0 = node0;
      ci = (int)pl.chans.size();
      pl.chans.push_back(cd);
      pl.chan_process.push_back(us.process);
      pl.chan_channel.push_back(us.channel);
      chan_index[key] = ci;
      pl.n_banks = std::max(pl.n_banks, (int)cd.n_banks);
    } else {
      ci = it->second;
    }
    SegDesc sd{};
    sd.channel = ci;
    sd.n_events = us.n_events;
    sd.x = us.x; sd.p = us.p; sd.factor = us.factors; sd.volume = us.volumes; sd.ok = us.oks;
    sd.align16 = ((reinterpret_cast<uintptr_t>(us.x) % 16) == 0 ? 1 : 0) |
                 ((reinterpret_cast<uintptr_t>(us.p) % 16) == 0 ? 2 : 0);
    const size_t k = pl.segs.size();
    sd.e1 = k < pl.beam_e1.size() ? pl.beam_e1[k] : nullptr;
    sd.e2 = k < pl.beam_e2.size() ? pl.beam_e2[k] : nullptr;
    sd.msq = k < pl.seg_msq.size() ? pl.seg_msq[k] : nullptr;
    pl.segs.push_back(sd);
  }
  if (pl.n_banks > 3) return set_error(GPU_PHS_E_TREE, "tree needs more than 3 save banks");
  // class of every channel
  pl.chan_class.assign(pl.chans.size(), 0);
  for (size_t c = 0; c < pl.chans.size(); c++) {
    const ChannelDesc& cd = pl.chans[c];
    bool small = cd.n_x <= 8;
    for (int i = 0; i < cd.n_nodes; i++) {
      if (i > 0 && pl.nodes[cd.node0 + i].push >= 0) small = false;
      if (pl.nodes[cd.node0 + i].friend_sign != 0) small = false;   // friend trees run in class 1
    }
    pl.chan_class[c] = small ? 0 : 1;
  }
  // tiles of each segment (segment order, for the plain launches), then the table in class
  // order, and within a class in segment order or deepest programs first
  std::vector<TileDesc> tiles;
  std::vector<int> tile_depth, tile_class;
  pl.seg_tile_begin.assign(pl.segs.size(), 0);
  pl.seg_tile_end.assign(pl.segs.size(), 0);
  for (size_t s = 0; s < pl.segs.size(); s++) {
    const int cl = pl.chan_class[pl.segs[s].channel];
    const int64_t tile = (int64_t)pl.lcs[cl].block * pl.lcs[cl].ept;
    for (int64_t e0 = 0; e0 < pl.segs[s].n_events; e0 += tile) {
      TileDesc td{};
      td.seg = (int32_t)s;
      td.e0 = e0;
      td.count = (int32_t)std::min<int64_t>(tile, pl.segs[s].n_events - e0);
      tiles.push_back(td);
      tile_depth.push_back(pl.chans[pl.segs[s].channel].n_nodes);
      tile_class.push_back(cl);
    }
  }
  if (tiles.size() > (size_t)0x7fffffff) return set_error(GPU_PHS_E_ARG, "too many tiles");
  std::vector<int> idx(tiles.size());
  std::iota(idx.begin(), idx.end(), 0);
  // each class grouped by tree shape (one specialised launch per shape); shapeless channels,
  // and shapes with fewer than merge_below events in the class, last (one generic launch)
  std::vector<int64_t> shape_events((size_t)kTileClasses * kNumShapes, 0);
  for (size_t t = 0; t < tiles.size(); t++) {
    const int sh = pl.chans[pl.segs[tiles[t].seg].channel].shape;
    if (sh >= 0) shape_events[(size_t)tile_class[t] * kNumShapes + sh] += tiles[t].count;
  }
  auto shape_key = [&](int t) {
    const int sh = pl.chans[pl.segs[tiles[t].seg].channel].shape;
    if (sh < 0 || shape_events[(size_t)tile_class[t] * kNumShapes + sh] < pl.merge_below) return kNumShapes;
    return sh;
  };
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    if (tile_class[a] != tile_class[b]) return tile_class[a] < tile_class[b];
    if (shape_key(a) != shape_key(b)) return shape_key(a) < shape_key(b);
    // order 1: heaviest programs first, so the tail of the launch is made of short tiles
    return pl.order == 1 && tile_depth[a] > tile_depth[b];
  });
  pl.tiles.resize(tiles.size());
  for (int c = 0; c < kTileClasses; c++) pl.class_begin[c] = pl.class_end[c] = 0;
  for (size_t i = 0; i < idx.size(); i++) {
    const TileDesc& td = tiles[idx[i]];
    pl.tiles[i] = td;
    const int cl = tile_class[idx[i]];
    if (pl.class_end[cl] == 0) pl.class_begin[cl] = (int)i;
    pl.class_end[cl] = (int)i + 1;
  }
  // plain launches: the table positions of each segment's tiles (contiguous per segment,
  // except under order 1, where a segment keeps its own range only if its tiles stay adjacent)
  {
    std::vector<int> pos(tiles.size());
    for (size_t i = 0; i < idx.size(); i++) pos[idx[i]] = (int)i;
    size_t t = 0;
    for (size_t s = 0; s < pl.segs.size(); s++) {
      int lo = 0x7fffffff, hi = 0;
      const size_t t0 = t;
      while (t < tiles.size() && tiles[t].seg == (int32_t)s) { lo = std::min(lo, pos[t]); hi = std::max(hi, pos[t] + 1); t++; }
      if (t == t0) lo = hi = 0;
      pl.seg_tile_begin[s] = lo;
      pl.seg_tile_end[s] = hi;
    }
  }

  // upload: one allocation, 256-byte aligned sections
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t o_tiles = 0, o_segs = al(pl.tiles.size() * sizeof(TileDesc));
  size_t o_chans = o_segs + al(pl.segs.size() * sizeof(SegDesc));
  size_t o_nodes = o_chans + al(pl.chans.size() * sizeof(ChannelDesc));
  size_t total = o_nodes + al(pl.nodes.size() * sizeof(NodeCmd)) + 256;
  drop_graph(pl);
  if (total > pl.dev_bytes) {
    if (pl.dev) (void)hipFree(pl.dev);
    pl.dev = nullptr;
    pl.dev_bytes = 0;
    hipError_t e = hipMalloc(&pl.dev, total);
    if (e != hipSuccess) return hip_error(e, "hipMalloc(plan tables)");
    pl.dev_bytes = total;
  }
  std::vector<char> host(total, 0);
  if (!pl.tiles.empty()) std::memcpy(host.data() + o_tiles, pl.tiles.data(), pl.tiles.size() * sizeof(TileDesc));
  if (!pl.segs.empty()) std::memcpy(host.data() + o_segs, pl.segs.data(), pl.segs.size() * sizeof(SegDesc));
  if (!pl.chans.empty()) std::memcpy(host.data() + o_chans, pl.chans.data(), pl.chans.size() * sizeof(ChannelDesc));
  if (!pl.nodes.empty()) std::memcpy(host.data() + o_nodes, pl.nodes.data(), pl.nodes.size() * sizeof(NodeCmd));
  hipError_t e = hipMemcpy(pl.dev, host.data(), total, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_error(e, "hipMemcpy(plan tables)");
  char* base = (char*)pl.dev;
  pl.dp.tiles = (const TileDesc*)(base + o_tiles);
  pl.dp.segs = (const SegDesc*)(base + o_segs);
  pl.dp.chans = (const ChannelDesc*)(base + o_chans);
  pl.dp.nodes = (const NodeCmd*)(base + o_nodes);
  pl.dp.n_tiles = (int)pl.tiles.size();
  // maxima over all channels (adjoint) and over each class (forward launches)
  auto maxima = [&](DevicePlan& d, int cl) {
    d.n_banks = 0; d.nx_max = d.nout_max = 0; d.root_banks_only = 1; d.has_friends = 0; d.per_event_beams = 0;
    for (size_t c = 0; c < pl.chans.size(); c++) {
      if (cl >= 0 && pl.chan_class[c] != cl) continue;
      const ChannelDesc& cd = pl.chans[c];
      d.n_banks = std::max(d.n_banks, (int)cd.n_banks);
      d.nx_max = std::max(d.nx_max, (int)cd.n_x);
      d.nout_max = std::max(d.nout_max, (int)cd.n_out);
      for (int i = 0; i < cd.n_nodes; i++) {
        const NodeCmd& nd = pl.nodes[cd.node0 + i];
        if (i > 0 && nd.push >= 0) d.root_banks_only = 0;
        // a save bank of the root's six boost parameters re-forms B_z R_z R_y, not a friend's matrix
        if (nd.friend_sign != 0) { d.has_friends = 1; if (i == 0) d.root_banks_only = 0; }
      }
    }
  };
  maxima(pl.dp, -1);
  for (int c = 0; c < kTileClasses; c++) {
    pl.dps[c] = pl.dp;
    maxima(pl.dps[c], c);
  }
  for (const SegDesc& sd : pl.segs)
    if (sd.e1) pl.dp.per_event_beams = pl.dps[pl.chan_class[sd.channel]].per_event_beams = 1;
  // each class by shape (with per-event beams, only the trees of up to four final-state
  // particles have specialised kernels; the others run the generic one)
  for (int c = 0; c < kTileClasses; c++) {
    std::vector<ShapeRun>& runs = pl.shape_runs[c];
    runs.clear();
    if (!pl.specialize) continue;
    const bool pev = pl.dps[c].per_event_beams != 0;
    for (int t = pl.class_begin[c]; t < pl.class_end[c]; t++) {
      int sh = pl.chans[pl.segs[pl.tiles[t].seg].channel].shape;
      if (sh >= 0 && shape_events[(size_t)c * kNumShapes + sh] < pl.merge_below) sh = -1;
      if (pev && sh >= kPevShapes) sh = -1;
      if (runs.empty() || runs.back().sid != sh) runs.push_back(ShapeRun{sh, t, t + 1});
      else runs.back().end = t + 1;
    }
  }
  // the launches, heaviest first, for spreading them over streams
  pl.jobs.clear();
  auto tile_bytes = [&](int t) {
    const ChannelDesc& cd = pl.chans[pl.segs[pl.tiles[t].seg].channel];
    return (double)pl.tiles[t].count * (8.0 * cd.n_x + 32.0 * cd.n_out + 17.0);
  };
  for (int c = 0; c < kTileClasses; c++) {
    if (pl.class_end[c] <= pl.class_begin[c]) continue;
    std::vector<ShapeRun> runs = pl.shape_runs[c];
    if (runs.empty()) runs.push_back(ShapeRun{-1, pl.class_begin[c], pl.class_end[c]});
    for (const ShapeRun& r : runs) {
      double b = 0;
      for (int t = r.begin; t < r.end; t++) b += tile_bytes(t);
      pl.jobs.push_back(PlanObj::Job{c, r.begin, r.end, r.sid, b});
    }
  }
  std::stable_sort(pl.jobs.begin(), pl.jobs.end(),
                   [](const PlanObj::Job& a, const PlanObj::Job& b) { return a.bytes > b.bytes; });
  return ok_status();
}

// Tiles [b, e) of class c: per shape run of class 0 the specialised kernel, else the generic one.
hipError_t launch_range(PlanObj& pl, int c, int b, int e, const LaunchConfig& lc, hipStream_t stream) {
  if (pl.shape_runs[c].empty()) return launch_forward(pl.dps[c], b, e, lc, stream);
  for (const ShapeRun& r : pl.shape_runs[c]) {
    const int lo = std::max(b, r.begin), hi = std::min(e, r.end);
    if (lo >= hi) continue;
    hipError_t err = r.sid >= 0 ? launch_forward_shape(pl.dps[c], r.sid, lo, hi, lc, stream)
                                : launch_forward(pl.dps[c], lo, hi, lc, stream);
    if (err != hipSuccess) return err;
  }
  return hipSuccess;
}

hipError_t launch_class(PlanObj& pl, int c, hipStream_t stream) {
  return launch_range(pl, c, pl.class_begin[c], pl.class_end[c], pl.lcs[c], stream);
}

// Both classes on `stream`'s timeline: class 0 on `stream`, class 1 (when both have tiles) on
// the side stream between a fork and a join event, so the two kernels can share the GPU.
// Every launch of the plan over n_streams streams: heaviest first onto the least loaded one.
hipError_t launch_jobs(PlanObj& pl, hipStream_t stream) {
  const int ns = std::min<int>(pl.n_streams, (int)pl.jobs.size());
  hipError_t e;
  if (!pl.fork_event && (e = hipEventCreateWithFlags(&pl.fork_event, hipEventDisableTiming)) != hipSuccess) return e;
  for (int k = 0; k < ns - 1; k++) {
    if (!pl.lane_streams[k] && (e = hipStreamCreateWithFlags(&pl.lane_streams[k], hipStreamNonBlocking)) != hipSuccess)
      return e;
    if (!pl.lane_join[k] && (e = hipEventCreateWithFlags(&pl.lane_join[k], hipEventDisableTiming)) != hipSuccess)
      return e;
  }
  if ((e = hipEventRecord(pl.fork_event, stream)) != hipSuccess) return e;
  for (int k = 0; k < ns - 1; k++)
    if ((e = hipStreamWaitEvent(pl.lane_streams[k], pl.fork_event, 0)) != hipSuccess) return e;
  double load[kMaxStreams] = {0, 0, 0, 0};
  for (const PlanObj::Job& j : pl.jobs) {
    int k = 0;
    for (int q = 1; q < ns; q++)
      if (load[q] < load[k]) k = q;
    load[k] += j.bytes;
    hipStream_t st = k == 0 ? stream : pl.lane_streams[k - 1];
    e = j.sid >= 0 ? launch_forward_shape(pl.dps[j.cls], j.sid, j.begin, j.end, pl.lcs[j.cls], st)
                   : launch_forward(pl.dps[j.cls], j.begin, j.end, pl.lcs[j.cls], st);
    if (e != hipSuccess) return e;
  }
  for (int k = 0; k < ns - 1; k++) {
    if ((e = hipEventRecord(pl.lane_join[k], pl.lane_streams[k])) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(stream, pl.lane_join[k], 0)) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_classes(PlanObj& pl, hipStream_t stream) {
  if (pl.n_streams > 1 && pl.jobs.size() > 1) return launch_jobs(pl, stream);
  const bool two = pl.class_end[0] > pl.class_begin[0] && pl.class_end[1] > pl.class_begin[1];
  if (!two || pl.serial_classes) {
    for (int c = 0; c < kTileClasses; c++) {
      hipError_t e = launch_class(pl, c, stream);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  hipError_t e;
  if (!pl.side_stream && (e = hipStreamCreateWithFlags(&pl.side_stream, hipStreamNonBlocking)) != hipSuccess) return e;
  if (!pl.fork_event && (e = hipEventCreateWithFlags(&pl.fork_event, hipEventDisableTiming)) != hipSuccess) return e;
  if (!pl.join_event && (e = hipEventCreateWithFlags(&pl.join_event, hipEventDisableTiming)) != hipSuccess) return e;
  if ((e = hipEventRecord(pl.fork_event, stream)) != hipSuccess) return e;
  if ((e = hipStreamWaitEvent(pl.side_stream, pl.fork_event, 0)) != hipSuccess) return e;
  if ((e = launch_class(pl, 1, pl.side_stream)) != hipSuccess) return e;
  if ((e = launch_class(pl, 0, stream)) != hipSuccess) return e;
  if ((e = hipEventRecord(pl.join_event, pl.side_stream)) != hipSuccess) return e;
  return hipStreamWaitEvent(stream, pl.join_event, 0);
}

int plan_run(PlanObj& pl, hipStream_t stream) {
  hipError_t e = launch_classes(pl, stream);
  if (e != hipSuccess) return hip_error(e, "forward launch");
  return ok_status();
}

int plan_run_plain(PlanObj& pl, hipStream_t stream) {
  for (size_t s = 0; s < pl.segs.size(); s++) {
    const int cl = pl.chan_class[pl.segs[s].channel];
    LaunchConfig lc = pl.lcs[cl];
    lc.grid = 0;
    hipError_t e = launch_range(pl, cl, pl.seg_tile_begin[s], pl.seg_tile_end[s], lc, stream);
    if (e != hipSuccess) return hip_error(e, "forward launch (plain)");
  }
  return ok_status();
}

}  // namespace gphs

using namespace gphs;

extern "C" {

int gpu_phs_last_error(void) { return t_err; }
const char* gpu_phs_last_error_message(void) { return t_msg.c_str(); }

int gpu_phs_process_create(int n_in, int n_out, int n_channels, const int* d1, const int* d2,
                           const int* has_children, const int* friends, const int* map_ids,
                           const double* map_masses, const double* map_widths, const double* flv_masses,
                           const double* flv_widths, double E1, double E2) {
  std::lock_guard<std::mutex> lk(api_mutex());
  if (n_in != 2 || n_out < 2 || n_out > kMaxOut || n_channels < 1)
    return set_error(GPU_PHS_E_ARG, "need n_in = 2, 2 <= n_out <= 12, n_channels >= 1");
  if (!d1 || !d2 || !has_children || !map_ids || !map_masses || !map_widths || !flv_masses)
    return set_error(GPU_PHS_E_ARG, "tree arrays must not be NULL");
  ProcessSpec sp;
  sp.n_in = n_in; sp.n_out = n_out; sp.n_channels = n_channels;
  size_t np = (size_t)n_channels * sp.n_prt(), npo = (size_t)n_channels * sp.n_prt_out();
  sp.d1.assign(d1, d1 + np);
  sp.d2.assign(d2, d2 + np);
  sp.has_children.assign(has_children, has_children + np);
  if (friends) sp.friends.assign(friends, friends + np);
  else sp.friends.assign(np, 0);
  sp.map_ids.assign(map_ids, map_ids + npo);
  sp.map_masses.assign(map_masses, map_masses + npo);
  sp.map_widths.assign(map_widths, map_widths + npo);
  sp.flv_masses.assign(flv_masses, flv_masses + n_in + n_out);
  if (flv_widths) sp.flv_widths.assign(flv_widths, flv_widths + n_in + n_out);
  else sp.flv_widths.assign(n_in + n_out, 0.0);
  sp.E1 = E1; sp.E2 = E2;
  auto po = std::make_unique<ProcessObj>();
  std::string err = compile_process(sp, po->cp);
  if (!err.empty()) return set_error(GPU_PHS_E_TREE, err);
  int h = register_process(std::move(po));
  ok_status();
  return h;
}

int gpu_phs_process_destroy(int process) {
  std::lock_guard<std::mutex> lk(api_mutex());
  if (drop_process(process) != GPU_PHS_OK) return set_error(GPU_PHS_E_HANDLE, "unknown process");
  return ok_status();
}

int gpu_phs_process_info(int process, int* n_out, int* n_x, int* n_channels) {
  std::lock_guard<std::mutex> lk(api_mutex());
  ProcessObj* po = find_process(process);
  if (!po) return set_error(GPU_PHS_E_HANDLE, "unknown process");
  if (n_out) *n_out = po->cp.spec.n_out;
  if (n_x) *n_x = po->cp.spec.n_x();
  if (n_channels) *n_channels = po->cp.spec.n_channels;
  return ok_status();
}

int gpu_phs_process_shapes(int process, int* shapes, int n) {
  std::lock_guard<std::mutex> lk(api_mutex());
  ProcessObj* po = find_process(process);
  if (!po) return set_error(GPU_PHS_E_HANDLE, "unknown process");
  if (!shapes || n < (int)po->cp.channels.size()) return set_error(GPU_PHS_E_ARG, "shapes: need one entry per channel");
  for (size_t c = 0; c < po->cp.channels.size(); c++) shapes[c] = po->cp.channels[c].shape;
  return ok_status();
}

int gpu_phs_process_trees(int process, int* d1, int* d2, int* has_children, int* map_ids, double* map_masses,
                          double* map_widths, double* flv_masses) {
  std::lock_guard<std::mutex> lk(api_mutex());
  ProcessObj* po = find_process(process);
  if (!po) return set_error(GPU_PHS_E_HANDLE, "unknown process");
  const ProcessSpec& sp = po->cp.spec;
  auto cp = [](auto* dst, const auto& v) { if (dst) std::copy(v.begin(), v.end(), dst); };
  cp(d1, sp.d1); cp(d2, sp.d2); cp(has_children, sp.has_children); cp(map_ids, sp.map_ids);
  cp(map_masses, sp.map_masses); cp(map_widths, sp.map_widths); cp(flv_masses, sp.flv_masses);
  return ok_status();
}

int gpu_phs_plan_create(const gpu_phs_segment* segments, int n_segments) {
  std::lock_guard<std::mutex> lk(api_mutex());
  if (n_segments < 1 || !segments) return set_error(GPU_PHS_E_ARG, "need at least one segment");
  int rc = require_device();
  if (rc) return rc;
  auto pl = std::make_unique<PlanObj>();
  pl->user.assign(segments, segments + n_segments);
  rc = plan_build(*pl);
  if (rc) return rc;
  // default workgroup: the largest whose staging tables fit the LDS of a workgroup
  for (int c = 0; c < kTileClasses; c++) {
    while (pl->lcs[c].block > 64 && forward_lds_bytes(pl->dps[c], pl->lcs[c]) > forward_lds_limit()) {
      pl->lcs[c].block -= 64;
      rc = plan_build(*pl);
      if (rc) return rc;
    }
  }
  int h = register_plan(std::move(pl));
  ok_status();
  return h;
}

int gpu_phs_plan_set_event_beams(int plan, int segment, const double* E1, const double* E2) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (segment < 0 || segment >= (int)pl->user.size()) return set_error(GPU_PHS_E_ARG, "set_event_beams: no such segment");
  if (!E1 != !E2) return set_error(GPU_PHS_E_ARG, "set_event_beams: give both E1 and E2, or neither");
  pl->beam_e1.resize(pl->user.size(), nullptr);
  pl->beam_e2.resize(pl->user.size(), nullptr);
  pl->beam_e1[segment] = E1;
  pl->beam_e2[segment] = E2;
  drop_graph(*pl);
  pl->adj_valid = false;
  pl->inv_segs.clear();
  return plan_build(*pl);
}

int gpu_phs_plan_set_msq_output(int plan, int segment, double* msq) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (segment < 0 || segment >= (int)pl->user.size()) return set_error(GPU_PHS_E_ARG, "set_msq_output: no such segment");
  pl->seg_msq.resize(pl->user.size(), nullptr);
  pl->seg_msq[segment] = msq;
  drop_graph(*pl);
  pl->adj_valid = false;
  pl->inv_segs.clear();
  return plan_build(*pl);
}

int gpu_phs_plan_set_specialize(int plan, int on) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  pl->specialize = on ? 1 : 0;
  drop_graph(*pl);
  return plan_build(*pl);
}

int gpu_phs_plan_destroy(int plan) {
  std::lock_guard<std::mutex> lk(api_mutex());
  if (drop_plan(plan) != GPU_PHS_OK) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  return ok_status();
}

int gpu_phs_plan_set_config(int plan, int block, int ept, int grid, int order, int waves_per_eu, int lds) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (block < 64 || block > 256 || block % 64 || ept < 1 || ept > 16 || grid < 0 || order < 0 || order > 1 ||
      !(waves_per_eu == 0 || (waves_per_eu >= 3 && waves_per_eu <= 5)) || lds < 0 || lds > 3)
    return set_error(GPU_PHS_E_ARG,
                     "block in {64,128,192,256}, 1 <= ept <= 16, grid >= 0, order in {0,1}, waves_per_eu in {0,3,4,5}, "
                     "lds in {0,1,2,3}");
  LaunchConfig want;
  want.block = block; want.ept = ept; want.grid = grid; want.waves_per_eu = waves_per_eu; want.lds = lds;
  for (int c = 0; c < kTileClasses; c++)
    if (pl->class_end[c] > pl->class_begin[c] && forward_lds_bytes(pl->dps[c], want) > forward_lds_limit())
      return set_error(GPU_PHS_E_ARG, "the staging tables of this block size (and LDS save banks) exceed the LDS a "
                                      "workgroup may use on this device; choose a smaller block");
  bool rebuild = order != pl->order;
  for (int c = 0; c < kTileClasses; c++) {
    const LaunchConfig& o = pl->lcs[c];
    if (block != o.block || ept != o.ept) rebuild = true;
    if (grid != o.grid || waves_per_eu != o.waves_per_eu || lds != o.lds) drop_graph(*pl);
    pl->lcs[c] = want;
  }
  pl->order = order;
  if (rebuild) return plan_build(*pl);
  return ok_status();
}

int gpu_phs_plan_set_class_config(int plan, int cls, int block, int ept, int grid, int waves_per_eu, int lds) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (cls < 0 || cls >= kTileClasses || block < 64 || block > 256 || block % 64 || ept < 1 || ept > 16 || grid < 0 ||
      !(waves_per_eu == 0 || (waves_per_eu >= 3 && waves_per_eu <= 5)) || lds < 0 || lds > 3)
    return set_error(GPU_PHS_E_ARG, "class in {0,1}, block in {64,128,192,256}, 1 <= ept <= 16, grid >= 0, "
                                    "waves_per_eu in {0,3,4,5}, lds in {0,1,2,3}");
  LaunchConfig want;
  want.block = block; want.ept = ept; want.grid = grid; want.waves_per_eu = waves_per_eu; want.lds = lds;
  if (pl->class_end[cls] > pl->class_begin[cls] && forward_lds_bytes(pl->dps[cls], want) > forward_lds_limit())
    return set_error(GPU_PHS_E_ARG, "the staging tables of this block size (and LDS save banks) exceed the LDS a "
                                    "workgroup may use on this device; choose a smaller block");
  const bool rebuild = block != pl->lcs[cls].block || ept != pl->lcs[cls].ept;
  pl->lcs[cls] = want;
  drop_graph(*pl);
  if (rebuild) return plan_build(*pl);
  return ok_status();
}

int gpu_phs_plan_get_class_config(int plan, int cls, int* block, int* ept, int* grid, int* waves_per_eu, int* lds) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (cls < 0 || cls >= kTileClasses) return set_error(GPU_PHS_E_ARG, "class in {0,1}");
  const LaunchConfig& lc = pl->lcs[cls];
  if (block) *block = lc.block;
  if (ept) *ept = lc.ept;
  if (grid) *grid = lc.grid;
  if (waves_per_eu) *waves_per_eu = lc.waves_per_eu;
  if (lds) *lds = lc.lds;
  return ok_status();
}

int gpu_phs_plan_set_schedule(int plan, int order, int serial) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (order < 0 || order > 1 || serial < 0 || serial > 1) return set_error(GPU_PHS_E_ARG, "order and serial in {0,1}");
  pl->serial_classes = serial;
  drop_graph(*pl);
  if (order != pl->order) {
    pl->order = order;
    return plan_build(*pl);
  }
  return ok_status();
}

int gpu_phs_plan_set_merge_below(int plan, int64_t events) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (events < 0) return set_error(GPU_PHS_E_ARG, "merge_below must not be negative");
  if (events == pl->merge_below) return ok_status();
  pl->merge_below = events;
  drop_graph(*pl);
  return plan_build(*pl);
}

int gpu_phs_plan_set_streams(int plan, int n_streams) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (n_streams < 1 || n_streams > kMaxStreams) return set_error(GPU_PHS_E_ARG, "n_streams in 1..4");
  pl->n_streams = n_streams;
  drop_graph(*pl);
  return ok_status();
}

int gpu_phs_plan_get_streams(int plan, int* n_streams) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (n_streams) *n_streams = pl->n_streams;
  return ok_status();
}

int gpu_phs_plan_get_merge_below(int plan, int64_t* events) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (events) *events = pl->merge_below;
  return ok_status();
}

int gpu_phs_plan_get_schedule(int plan, int* order, int* serial) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (order) *order = pl->order;
  if (serial) *serial = pl->serial_classes;
  return ok_status();
}

int gpu_phs_plan_get_config(int plan, int* block, int* ept, int* grid, int* order, int* waves_per_eu, int* lds) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  const LaunchConfig& lc = pl->lcs[pl->class_end[0] > pl->class_begin[0] ? 0 : 1];
  if (block) *block = lc.block;
  if (ept) *ept = lc.ept;
  if (grid) *grid = lc.grid;
  if (order) *order = pl->order;
  if (waves_per_eu) *waves_per_eu = lc.waves_per_eu;
  if (lds) *lds = lc.lds;
  return ok_status();
}

int gpu_phs_plan_info(int plan, int* n_tiles, int* n_segments, int* n_banks) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (n_tiles) *n_tiles = pl->dp.n_tiles;
  if (n_segments) *n_segments = (int)pl->segs.size();
  if (n_banks) *n_banks = pl->n_banks;
  return ok_status();
}

int gpu_phs_plan_run(int plan, void* stream) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  return plan_run(*pl, (hipStream_t)stream);
}

int gpu_phs_plan_run_plain(int plan, void* stream) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  return plan_run_plain(*pl, (hipStream_t)stream);
}

int gpu_phs_plan_run_graph(int plan, void* stream) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  hipError_t e;
  if (!pl->graph_exec) {
    if (!pl->capture_stream) {
      e = hipStreamCreateWithFlags(&pl->capture_stream, hipStreamNonBlocking);
      if (e != hipSuccess) return hip_error(e, "hipStreamCreate");
    }
    e = hipStreamBeginCapture(pl->capture_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_error(e, "hipStreamBeginCapture");
    hipError_t le = launch_classes(*pl, pl->capture_stream);
    e = hipStreamEndCapture(pl->capture_stream, &pl->graph);
    if (le != hipSuccess) return hip_error(le, "forward launch (capture)");
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&pl->graph_exec, pl->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return hip_error(e, "hipGraphInstantiate");
  }
  e = hipGraphLaunch(pl->graph_exec, (hipStream_t)stream);
  if (e != hipSuccess) return hip_error(e, "hipGraphLaunch");
  return ok_status();
}

int gpu_phs_plan_autotune(int plan, void* stream_v, int reps, char* report, int report_len) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  hipStream_t stream = (hipStream_t)stream_v;
  if (reps < 1) reps = 3;
  int dev = 0, n_cu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  if (n_cu < 1) n_cu = 256;
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
    return set_error(GPU_PHS_E_HIP, "hipEventCreate");
  // time one class's launch (cls >= 0) or the whole plan (cls < 0) under the current tables
  // Warm-up before a timing: the launches themselves, back to back, for at least 20 ms of GPU
  // time (at most 256 of them) -- after a rebuild of the tables the GPU has idled, and a
  // single warm-up launch left the next timings of a 2 ms step 20-30% slow.
  auto warm_up = [&](const std::function<int()>& run) -> int {
    float total = 0;
    for (int w = 0; w < 256 && total < 20.0f; w++) {
      (void)hipEventRecord(a, stream);
      int rc = run();
      if (rc) return rc;
      (void)hipEventRecord(b, stream);
      if (hipEventSynchronize(b) != hipSuccess) return set_error(GPU_PHS_E_HIP, "hipEventSynchronize");
      float x = 0;
      (void)hipEventElapsedTime(&x, a, b);
      total += x;
    }
    return GPU_PHS_OK;
  };
  auto time_now = [&](int cls, float* ms) -> int {
    std::function<int()> run = [&]() -> int {
      if (cls < 0) return plan_run(*pl, stream);
      hipError_t e = launch_class(*pl, cls, stream);
      return e == hipSuccess ? GPU_PHS_OK : hip_error(e, "forward launch");
    };
    int rc = warm_up(run);
    if (rc) return rc;
    std::vector<float> t;
    for (int r = 0; r < reps; r++) {
      (void)hipEventRecord(a, stream);
      rc = run();
      if (rc) return rc;
      (void)hipEventRecord(b, stream);
      if (hipEventSynchronize(b) != hipSuccess) return set_error(GPU_PHS_E_HIP, "hipEventSynchronize");
      float x = 0;
      (void)hipEventElapsedTime(&x, a, b);
      t.push_back(x);
    }
    std::sort(t.begin(), t.end());
    *ms = t[t.size() / 2];
    return GPU_PHS_OK;
  };
  struct Cfg { int cls; LaunchConfig lc; int order; float ms; };
  std::vector<Cfg> tried;
  int rc = GPU_PHS_OK;
  for (int cls = 0; cls < kTileClasses && !rc; cls++) {
    if (pl->class_end[cls] <= pl->class_begin[cls]) continue;
    Cfg best{cls, pl->lcs[cls], pl->order, 1e30f};
    auto consider = [&](LaunchConfig lc) -> int {
      if (forward_lds_bytes(pl->dps[cls], lc) > forward_lds_limit()) return GPU_PHS_OK;   // does not fit
      const bool rebuild = lc.block != pl->lcs[cls].block || lc.ept != pl->lcs[cls].ept;
      pl->lcs[cls] = lc;
      drop_graph(*pl);
      if (rebuild) {
        int r = plan_build(*pl);
        if (r) return r;
      }
      float ms = 0;
      int r = time_now(cls, &ms);
      if (r) return r;
      tried.push_back({cls, lc, pl->order, ms});
      // a new shape has to win by 2%: timings of one shape vary by about that much, and a
      // switch on noise would leave a worse shape in place
      if (ms < 0.98f * best.ms) best = {cls, lc, pl->order, ms};
      return GPU_PHS_OK;
    };
    const bool banks = pl->dps[cls].n_banks > 0;
    // the shape-specialised kernels keep randoms and banks in registers: where only they run,
    // the placement flags change nothing; those of up to four particles also have a budget for
    // 5 waves per SIMD, the larger ones only the compiler's
    bool placement = pl->shape_runs[cls].empty(), small = false;
    for (const ShapeRun& r : pl->shape_runs[cls]) {
      placement = placement || r.sid < 0;
      small = small || (r.sid >= 0 && kShapes[r.sid].n_out <= 4);
    }
    const int wpe_max = placement ? 4 : (small ? 5 : 0);
    // stage 1: register budget (occupancy target), where deferred boosts wait (lds bit 0) and
    // where the randoms are kept (lds bit 1)
    for (int wpe : {0, 3, 4, 5})
      for (int lds : {0, 1, 2, 3})
        if (!rc && wpe <= wpe_max && (!(lds & 1) || banks) && (placement || lds == 0)) {
          LaunchConfig lc = best.lc;
          lc.waves_per_eu = wpe; lc.lds = lds; lc.grid = 0;
          rc = consider(lc);
        }
    // stage 2: workgroup size x events per thread, with the randoms where stage 1 put them
    // and in the other place (registers or LDS: which wins depends on the shape)
    const LaunchConfig base = best.lc;
    for (int block : {64, 128, 256})
      for (int ept : {1, 2, 4})
        for (int xl : {0, 2})
          if (!rc && (placement || xl == 0)) {
            LaunchConfig lc = base;
            lc.block = block; lc.ept = ept; lc.lds = base.lds ^ xl;
            rc = consider(lc);
          }
    // stage 1 again at the chosen workgroup shape, if that changed: the best register budget
    // and placement depend on it
    if (best.lc.block != base.block || best.lc.ept != base.ept) {
      const LaunchConfig shape = best.lc;
      for (int wpe : {0, 3, 4, 5})
        for (int lds : {0, 1, 2, 3})
          if (!rc && wpe <= wpe_max && (!(lds & 1) || banks) && (placement || lds == 0) &&
              !(wpe == shape.waves_per_eu && lds == shape.lds)) {
            LaunchConfig lc = shape;
            lc.waves_per_eu = wpe; lc.lds = lds;
            rc = consider(lc);
          }
    }
    // stage 3: persistent grids around the best shape
    const LaunchConfig b2 = best.lc;
    for (int k : {2, 4, 8})
      if (!rc) {
        LaunchConfig lc = b2;
        lc.grid = n_cu * k * (256 / b2.block);
        rc = consider(lc);
      }
    if (!rc) {
      pl->lcs[cls] = best.lc;
      drop_graph(*pl);
      rc = plan_build(*pl);
    }
  }
  // The whole-plan stages time their candidates in interleaved rounds (each candidate once per
  // round, the median over the rounds), so that a clock still settling during a stage favours
  // none of them; a later candidate has to beat the best earlier one by 2%.
  auto stage = [&](int n, const std::function<int(int)>& apply, float* best_ms) -> int {
    std::vector<std::vector<float>> t(n);
    const int rounds = std::max(3, reps);
    for (int r = 0; r < rounds && !rc; r++)
      for (int i = 0; i < n && !rc; i++) {
        rc = apply(i);
        if (rc) break;
        rc = warm_up([&]() { return plan_run(*pl, stream); });
        if (rc) break;
        (void)hipEventRecord(a, stream);
        rc = plan_run(*pl, stream);
        if (rc) break;
        (void)hipEventRecord(b, stream);
        if (hipEventSynchronize(b) != hipSuccess) { rc = set_error(GPU_PHS_E_HIP, "hipEventSynchronize"); break; }
        float x = 0;
        (void)hipEventElapsedTime(&x, a, b);
        t[i].push_back(x);
      }
    if (rc) return 0;
    int best = -1;
    float bm = 1e30f;
    for (int i = 0; i < n; i++) {
      std::sort(t[i].begin(), t[i].end());
      const float m = t[i][t[i].size() / 2];
      if (m < 0.98f * bm) { best = i; bm = m; }
    }
    *best_ms = bm;
    return best;
  };
  // tile order, and whether two classes share the GPU or run one after the other
  pl->n_streams = 1;
  const bool two = pl->class_end[0] > pl->class_begin[0] && pl->class_end[1] > pl->class_begin[1];
  std::vector<std::pair<int, int>> scheds;
  for (int order : {0, 1})
    for (int serial : {0, 1})
      if (!serial || two) scheds.push_back({order, serial});
  float best_order_ms = 0;
  int k = stage((int)scheds.size(), [&](int i) {
    pl->serial_classes = scheds[i].second;
    drop_graph(*pl);
    if (pl->order == scheds[i].first) return (int)GPU_PHS_OK;
    pl->order = scheds[i].first;
    return plan_build(*pl);
  }, &best_order_ms);
  const int best_order = rc ? pl->order : scheds[k].first, best_serial = rc ? pl->serial_classes : scheds[k].second;
  // then the size below which a tree shape joins its class's generic launch
  if (!rc) {
    pl->order = best_order;
    pl->serial_classes = best_serial;
    drop_graph(*pl);
    rc = plan_build(*pl);
  }
  int64_t best_merge = pl->merge_below;
  float best_merge_ms = 0;
  if (!rc && pl->specialize) {
    const int64_t merges[] = {kMergeBelowDefault, 0, (int64_t)1 << 16, (int64_t)1 << 20};
    k = stage(4, [&](int i) {
      pl->merge_below = merges[i];
      drop_graph(*pl);
      return plan_build(*pl);
    }, &best_merge_ms);
    if (!rc) best_merge = merges[k];
  }
  if (!rc) {
    pl->merge_below = best_merge;
    drop_graph(*pl);
    rc = plan_build(*pl);
  }
  // then over how many streams the launches spread (1: the classes as chosen above)
  int best_streams = 1;
  float best_streams_ms = 0;
  if (!rc && pl->jobs.size() > 1) {
    k = stage(kMaxStreams, [&](int i) {
      pl->n_streams = i + 1;
      drop_graph(*pl);
      return (int)GPU_PHS_OK;
    }, &best_streams_ms);
    if (!rc) best_streams = k + 1;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc) return rc;
  pl->n_streams = best_streams;
  drop_graph(*pl);
  if (report && report_len > 0) {
    std::string s;
    char buf[128];
    for (int c = 0; c < kTileClasses; c++) {
      if (pl->class_end[c] <= pl->class_begin[c]) continue;
      const LaunchConfig& lc = pl->lcs[c];
      snprintf(buf, sizeof buf, "class %d: block=%d ept=%d grid=%d wpe=%d lds=%d; ", c, lc.block, lc.ept, lc.grid,
               lc.waves_per_eu, lc.lds);
      s += buf;
    }
    snprintf(buf, sizeof buf, "order=%d serial=%d %.4f ms; merge_below=%lld %.4f ms; streams=%d %.4f ms;", best_order,
             best_serial, best_order_ms, (long long)best_merge, best_merge_ms, best_streams, best_streams_ms);
    s += buf;
    for (const auto& c : tried) {
      snprintf(buf, sizeof buf, " c%d:%d/%d/%d/%d/%d:%.4f", c.cls, c.lc.block, c.lc.ept, c.lc.grid, c.lc.waves_per_eu,
               c.lc.lds, c.ms);
      s += buf;
    }
    snprintf(report, report_len, "%s", s.c_str());
  }
  return ok_status();
}

}  // extern "C"
