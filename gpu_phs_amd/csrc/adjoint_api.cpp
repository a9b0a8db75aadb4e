// C ABI of the adjoint (include/gpu_phs.h; no counterpart in the reference).
#include <algorithm>
#include <cstring>

#include "phs_state.h"

namespace gphs {
namespace {

// Device tables for the adjoint of a plan: the plan's tiles, channels and programs, a segment
// table carrying the gradient buffers, the per-tile partials and the tiles of each channel.
int adjoint_prepare(PlanObj& pl, const gpu_phs_adjoint_segment* adj, int write_forward) {
  const int n_seg = (int)pl.segs.size(), n_ch = (int)pl.chans.size(), n_t = pl.dp.n_tiles;

  int max_nodes = 0, max_banks = 0;
  for (const auto& c : pl.chans) {
    max_nodes = std::max(max_nodes, (int)c.n_nodes);
    max_banks = std::max(max_banks, (int)c.n_banks);
  }
  AdjLayout lay;
  // only the root defers (and the randoms fit the register row): the compact stash; it keeps
  // eight one-wavefront workgroups of the 2->4 trees within a CU's LDS
  int max_nx_all = 0;
  for (const auto& c : pl.chans) max_nx_all = std::max(max_nx_all, (int)c.n_x);
  lay.compact = pl.dp.root_banks_only && max_nx_all <= 8 && !pl.dp.has_friends ? 1 : 0;
  lay.off_in = 6 * max_nodes;
  lay.off_st = lay.off_in + 2 * max_nodes;
  lay.off_acc = lay.off_st + (lay.compact ? kAdjCompactStashWords : kAdjStashWords) * max_banks;
  // programs of up to 4 final-state particles keep their accumulators in registers (not those
  // of plans with friends, which run the generic kernel)
  int max_nx = 0;
  for (const auto& c : pl.chans) max_nx = std::max(max_nx, (int)c.n_x);
  lay.words = lay.off_acc + (max_nx <= 8 && !pl.dp.has_friends ? 0 : 2 * max_nodes);
  if ((size_t)lay.words * 64 * sizeof(double) > forward_lds_limit())
    return set_error(GPU_PHS_E_TREE, "adjoint: trees too large for the LDS tape");
  std::vector<SegDesc> segs = pl.segs;
  for (int s = 0; s < n_seg; s++) {
    SegDesc& sd = segs[s];
    if (pl.segs[s].n_events > 0 && !adj[s].g_x) return set_error(GPU_PHS_E_ARG, "adjoint: g_x must not be NULL");
    sd.g_p = adj[s].g_p;
    sd.g_f = adj[s].g_f;
    sd.g_x = adj[s].g_x;
    sd.loss = adj[s].loss;
    if (!write_forward) { sd.p = nullptr; sd.factor = nullptr; sd.volume = nullptr; sd.ok = nullptr; }
  }
  // the same buffers as the previous call: the device tables are current
  if (pl.adj_valid && segs.size() == pl.adj_segs.size() &&
      std::memcmp(segs.data(), pl.adj_segs.data(), segs.size() * sizeof(SegDesc)) == 0)
    return GPU_PHS_OK;
  pl.adj_valid = false;
  pl.adj_lay = lay;
  pl.adj_part_stride = 2 * max_nodes;
  pl.adj_segs = std::move(segs);
  // one wavefront per 64 events of a tile (tile sizes of the classes that have tiles)
  int tile_max = 1;
  for (int c = 0; c < kTileClasses; c++)
    if (pl.class_end[c] > pl.class_begin[c]) tile_max = std::max(tile_max, pl.lcs[c].block * pl.lcs[c].ept);
  pl.adj_sub = (tile_max + 63) / 64;
  // tiles of every channel, in tile order
  std::vector<int> off(n_ch + 1, 0), list(n_t);
  for (const auto& t : pl.tiles) off[pl.segs[t.seg].channel + 1]++;
  for (int c = 0; c < n_ch; c++) off[c + 1] += off[c];
  std::vector<int> fill(off.begin(), off.end() - 1);
  for (int t = 0; t < n_t; t++) list[fill[pl.segs[pl.tiles[t].seg].channel]++] = t;

  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_segs = al(n_seg * sizeof(SegDesc));
  const size_t b_part = al((size_t)std::max(n_t, 1) * pl.adj_sub * pl.adj_part_stride * sizeof(double));
  const size_t b_list = al(std::max(n_t, 1) * sizeof(int)), b_off = al((n_ch + 1) * sizeof(int));
  const size_t b_ranges = al((size_t)std::max(n_ch, 1) * kAdjRanges * pl.adj_part_stride * sizeof(double));
  const size_t b_done = al((size_t)std::max(n_ch, 1) * sizeof(unsigned int));
  const size_t total = b_segs + b_part + b_list + b_off + b_ranges + b_done;
  if (total > pl.adj_bytes) {
    if (pl.adj_dev) (void)hipFree(pl.adj_dev);
    pl.adj_dev = nullptr;
    pl.adj_bytes = 0;
    hipError_t e = hipMalloc(&pl.adj_dev, total);
    if (e != hipSuccess) return hip_error(e, "hipMalloc(adjoint tables)");
    pl.adj_bytes = total;
  }
  char* base = (char*)pl.adj_dev;
  hipError_t e;
  if ((e = hipMemcpy(base, pl.adj_segs.data(), n_seg * sizeof(SegDesc), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(base + b_segs + b_part, list.data(), n_t * sizeof(int), hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(base + b_segs + b_part + b_list, off.data(), (n_ch + 1) * sizeof(int), hipMemcpyHostToDevice)) !=
          hipSuccess ||
      (e = hipMemset(base + b_segs + b_part + b_list + b_off + b_ranges, 0, b_done)) != hipSuccess)
    return hip_error(e, "hipMemcpy(adjoint tables)");
  pl.adj_dp = pl.dp;
  pl.adj_dp.segs = (const SegDesc*)base;
  pl.adj_part = (double*)(base + b_segs);
  pl.adj_chan_tiles = (int*)(base + b_segs + b_part);
  pl.adj_chan_off = (int*)(base + b_segs + b_part + b_list);
  pl.adj_ranges = (double*)(base + b_segs + b_part + b_list + b_off);
  pl.adj_done = (unsigned int*)(base + b_segs + b_part + b_list + b_off + b_ranges);
  pl.adj_valid = true;
  return GPU_PHS_OK;
}

}  // namespace
}  // namespace gphs

using namespace gphs;

extern "C" {

int gpu_phs_plan_n_params(int plan) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  ok_status();
  return 2 * (int)pl->nodes.size();
}

int gpu_phs_plan_param_map(int plan, int* process, int* channel, int* bincode, int n) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (n < (int)pl->nodes.size()) return set_error(GPU_PHS_E_ARG, "param_map: arrays shorter than the node count");
  for (size_t c = 0; c < pl->chans.size(); c++) {
    ProcessObj* po = find_process(pl->chan_process[c]);
    if (!po) return set_error(GPU_PHS_E_HANDLE, "param_map: process destroyed");
    const ChannelDesc& src = po->cp.channels[pl->chan_channel[c]];
    for (int v = 0; v < pl->chans[c].n_nodes; v++) {
      const int slot = pl->chans[c].param0 + v;
      if (process) process[slot] = pl->chan_process[c];
      if (channel) channel[slot] = pl->chan_channel[c];
      if (bincode) bincode[slot] = po->cp.node_branch[src.node0 + v] + 1;
    }
  }
  return ok_status();
}

int gpu_phs_plan_adjoint(int plan, const gpu_phs_adjoint_segment* adj, int loss_mode, int write_forward, double* gparam,
                         void* stream) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  if (!adj || !gparam || loss_mode < 0 || loss_mode > 1)
    return set_error(GPU_PHS_E_ARG, "adjoint: need adjoint segments, gparam and loss_mode in {0, 1}");
  for (const auto& c : pl->chans)
    if (c.bg_lab != 0 && loss_mode == 1)
      return set_error(GPU_PHS_E_ARG, "adjoint: the built-in loss assumes symmetric beams");
  if (pl->dp.per_event_beams && loss_mode == 1)
    return set_error(GPU_PHS_E_ARG, "adjoint: the built-in loss assumes symmetric beams; with per-event beam energies "
                                    "give the cotangents (loss_mode 0)");
  int rc = adjoint_prepare(*pl, adj, write_forward);
  if (rc) return rc;
  // the tile table in runs of one tree shape: a shape of up to four final-state particles
  // takes the adjoint specialised for it, and consecutive runs of such shapes one launch of the
  // multi-shape adjoint (so the shapes the forward pass folds into its generic launch, those
  // with fewer than merge_below events, keep their specialised adjoint); the others the generic
  // adjoint.  One workgroup per 64-event unit: a persistent grid measured slower (its
  // workgroups do not all fit the LDS at once), and units without events would cost a
  // workgroup each.
  std::vector<TileRun> runs;
  for (int c = 0; c < kTileClasses; c++) {
    for (int t = pl->class_begin[c]; t < pl->class_end[c]; t++) {
      const int sh = pl->chans[pl->segs[pl->tiles[t].seg].channel].shape;
      const int sid = pl->specialize && sh >= 0 && sh < kAdjShapes ? sh : -1;
      if (runs.empty() || runs.back().sid != sid || runs.back().end != t) runs.push_back(TileRun{sid, t, t + 1});
      else runs.back().end = t + 1;
    }
  }
  hipError_t e = launch_adjoint(pl->adj_dp, pl->adj_lay, runs.data(), (int)runs.size(), loss_mode, pl->adj_part,
                                pl->adj_part_stride, pl->adj_sub, pl->adj_ranges, pl->adj_done, pl->adj_chan_tiles,
                                pl->adj_chan_off, (int)pl->chans.size(), gparam, (hipStream_t)stream);
  if (e != hipSuccess) return hip_error(e, "adjoint launch");
  return ok_status();
}

}  // extern "C"
