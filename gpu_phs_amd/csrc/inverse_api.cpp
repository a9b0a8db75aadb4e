// C ABI of the inverse mapping (include/gpu_phs.h): all-channel factors and randoms of the
// events of a plan (reference: the do_inverse_mapping branch of gen_phs_from_x_gpu,
// phs.cu:1491-1554).
#include <algorithm>
#include <cstring>
#include <tuple>

#include "phs_state.h"

using namespace gphs;

namespace {

// Device tables of the inverse: per-segment (generating channel, first row), then the
// process's channels and programs.  Re-uploaded only when the rows change.
int inverse_prepare(PlanObj& pl, const int64_t* seg_row0, int64_t n_rows) {
  if (pl.segs.empty()) return set_error(GPU_PHS_E_ARG, "inverse: the plan has no segments");
  const int proc = pl.chan_process.empty() ? -1 : pl.chan_process[0];
  for (int p : pl.chan_process)
    if (p != proc) return set_error(GPU_PHS_E_ARG, "inverse: all segments must belong to one process");
  ProcessObj* po = find_process(proc);
  if (!po) return set_error(GPU_PHS_E_HANDLE, "inverse: the plan's process was destroyed");
  const CompiledProcess& cp = po->cp;
  std::vector<InvSeg> is(pl.segs.size());
  for (size_t s = 0; s < pl.segs.size(); s++) {
    const int64_t r0 = seg_row0[s], n = pl.segs[s].n_events;
    if (r0 < 0 || r0 + n > n_rows) return set_error(GPU_PHS_E_ARG, "inverse: segment rows outside [0, n_rows)");
    is[s].own = pl.chan_channel[pl.segs[s].channel];
    is[s].pad = 0;
    is[s].row0 = r0;
  }
  const bool shared = !cp.has_friends && pl.specialize;
  if (pl.inv_dev && pl.inv_process == proc && pl.inv.per_event_beams == pl.dp.per_event_beams &&
      pl.inv.shared_roots == (shared ? 1 : 0) && is.size() == pl.inv_segs.size() &&
      std::memcmp(is.data(), pl.inv_segs.data(), is.size() * sizeof(InvSeg)) == 0)
    return ok_status();
  const int n_ch = (int)cp.channels.size();
  // channel order: by the root's daughters (final-state particles and sum program of each), so
  // that channels sharing their root's evaluation are adjacent; channel index within a group
  auto root_key = [&](int c) {
    const NodeCmd& r = cp.nodes[cp.channels[c].node0];
    return std::make_tuple(r.k[0].leaves, r.k[0].sum_len, r.k[0].sum_code, r.k[1].leaves, r.k[1].sum_len,
                           r.k[1].sum_code);
  };
  std::vector<int> order(n_ch);
  for (int c = 0; c < n_ch; c++) order[c] = c;
  if (shared) std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return root_key(a) < root_key(b); });
  std::vector<InvStep> steps(n_ch);
  for (int k = 0; k < n_ch; k++) {
    steps[k].channel = order[k];
    steps[k].new_group = k == 0 || root_key(order[k]) != root_key(order[k - 1]) ? 1 : 0;
  }
  // save-bank levels of the subtrees deferred below the root (the root's own deferred daughter
  // is re-formed from the shared root)
  int rest_banks = 0;
  for (const ChannelDesc& c : cp.channels) {
    int depth = 0;
    for (int i = 1; i < c.n_nodes; i++) {
      const NodeCmd& nd = cp.nodes[c.node0 + i];
      if (nd.push >= 0) rest_banks = std::max(rest_banks, ++depth);
      if (nd.desc < 0 && depth > 0) depth--;
    }
  }
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t o_chans = al(is.size() * sizeof(InvSeg));
  const size_t o_nodes = o_chans + al(cp.channels.size() * sizeof(ChannelDesc));
  const size_t o_steps = o_nodes + al(cp.nodes.size() * sizeof(NodeCmd));
  const size_t total = o_steps + al(steps.size() * sizeof(InvStep));
  if (total > pl.inv_bytes) {
    if (pl.inv_dev) (void)hipFree(pl.inv_dev);
    pl.inv_dev = nullptr;
    pl.inv_bytes = 0;
    hipError_t e = hipMalloc(&pl.inv_dev, total);
    if (e != hipSuccess) return hip_error(e, "hipMalloc(inverse tables)");
    pl.inv_bytes = total;
  }
  std::vector<char> host(total, 0);
  std::memcpy(host.data(), is.data(), is.size() * sizeof(InvSeg));
  std::memcpy(host.data() + o_chans, cp.channels.data(), cp.channels.size() * sizeof(ChannelDesc));
  std::memcpy(host.data() + o_nodes, cp.nodes.data(), cp.nodes.size() * sizeof(NodeCmd));
  std::memcpy(host.data() + o_steps, steps.data(), steps.size() * sizeof(InvStep));
  hipError_t e = hipMemcpy(pl.inv_dev, host.data(), total, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_error(e, "hipMemcpy(inverse tables)");
  char* base = (char*)pl.inv_dev;
  pl.inv.isegs = (const InvSeg*)base;
  pl.inv.chans = (const ChannelDesc*)(base + o_chans);
  pl.inv.nodes = (const NodeCmd*)(base + o_nodes);
  pl.inv.steps = (const InvStep*)(base + o_steps);
  pl.inv.n_ch = n_ch;
  pl.inv.n_out = cp.spec.n_out;
  pl.inv.n_x = cp.spec.n_x();
  pl.inv.n_banks = cp.max_banks;
  int root_bank = cp.max_banks <= 1 ? 1 : 0;
  for (const ChannelDesc& c : cp.channels)
    for (int i = 0; i < c.n_nodes; i++) {
      const NodeCmd& nd = cp.nodes[c.node0 + i];
      if ((i > 0 && nd.push >= 0) || (i == 0 && nd.friend_sign != 0)) root_bank = 0;
    }
  pl.inv.root_bank = root_bank;
  pl.inv.has_friends = cp.has_friends ? 1 : 0;
  pl.inv.per_event_beams = pl.dp.per_event_beams;
  pl.inv.shared_roots = shared ? 1 : 0;
  pl.inv.rest_banks = rest_banks;
  pl.inv.generic = 0;
  for (const ChannelDesc& c : cp.channels)
    if (c.shape < 0 || c.shape >= kInvShapes) pl.inv.generic = 1;
  pl.inv_segs = std::move(is);
  pl.inv_process = proc;
  return ok_status();
}

}  // namespace

int gphs::plan_inverse(PlanObj& pl, const int64_t* seg_row0, int64_t n_rows, double* factors, double* x_out, int block,
                       hipStream_t stream) {
  if (!seg_row0 || n_rows <= 0 || !factors || !x_out)
    return set_error(GPU_PHS_E_ARG, "inverse: need seg_row0, n_rows > 0, factors and x_out");
  if (block != 0 && (block < 64 || block > 256 || block % 64))
    return set_error(GPU_PHS_E_ARG, "inverse: block must be 64, 128, 192 or 256");
  int rc = inverse_prepare(pl, seg_row0, n_rows);
  if (rc) return rc;
  // the default workgroup: 256 threads for processes of a few channels (the 3 c1 channels: 1.72
  // against 1.79 ms for 1e7 events at 128), 128 for many (the 120 c2 channels: 7.65 against
  // 7.70-7.83 ms for 1e6 events at 256), each where its LDS tables fit
  if (block == 0) {
    block = pl.inv.n_ch <= 16 ? 256 : 128;
    while (block > 64 && inverse_lds_bytes(pl.inv, block) > forward_lds_limit()) block -= 64;
  }
  if (inverse_lds_bytes(pl.inv, block) > forward_lds_limit())
    return set_error(GPU_PHS_E_ARG, "inverse: the trees need more LDS than a workgroup of this size may use");
  hipError_t e = launch_inverse(pl.inv, pl.dp.tiles, pl.dp.n_tiles, pl.dp.segs, block, 0, factors, x_out, n_rows, stream);
  if (e != hipSuccess) return hip_error(e, "inverse launch");
  return ok_status();
}

extern "C" {

int gpu_phs_plan_inverse(int plan, const int64_t* seg_row0, int64_t n_rows, double* factors, double* x_out,
                         int block, void* stream) {
  std::lock_guard<std::mutex> lk(api_mutex());
  PlanObj* pl = find_plan(plan);
  if (!pl) return set_error(GPU_PHS_E_HANDLE, "unknown plan");
  return plan_inverse(*pl, seg_row0, n_rows, factors, x_out, block, (hipStream_t)stream);
}

}  // extern "C"
