// Allocation-free forward pass of one channel (gpu_phs_generate_noalloc, include/gpu_phs.h).
//
// The batched plans (gpu_phs_plan_create / gpu_phs_plan_run) keep their tile, segment, channel
// and program tables in a buffer they allocate.  Callers that cannot allocate -- code under
// HIP graph capture, tight serving loops, harnesses that forbid it -- get this entry point
// instead: the same forward kernels, with the tables of the call in `scratch`, a fixed-size
// static device array that every GPU holds for the life of the process.  The tables reach it
// through the kernel arguments of two small staging kernels (the segment, channel and program
// records in chunks; the tile table computed from the event count), queued on the caller's
// stream before the forward launch, so no host buffer has to outlive the call and nothing is
// copied from host memory.  Every call on a GPU shares that GPU's array and overwrites it:
// calls that overlap on one GPU (different streams without ordering between them) race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>

#include "phs_shapes.h"
#include "phs_state.h"

namespace gphs {

namespace {

// Layout of `scratch`: segment, channel and program records in the first kTableBytes, then the
// tile table.
constexpr size_t kScratchBytes = (size_t)1 << 20;
constexpr size_t kTableBytes = 16 * 1024;
constexpr int64_t kMaxTiles = (int64_t)((kScratchBytes - kTableBytes) / sizeof(TileDesc));

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each device (hipGetSymbolAddress, once per device).
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};

// Fixed launch shapes (there is no autotuning here): the trees of up to four final-state
// particles at 256 threads and 4 events per thread, the others at 64 threads.
constexpr int kBlockSmall = 256, kBlockLarge = 64, kEpt = 4;

// A chunk of the records, by value in the kernel arguments (well under the 4 KiB limit).
constexpr int kChunkWords = 384;
struct StageChunk {
  uint32_t offset_words;   // destination in `scratch`, in 8-byte words
  uint32_t n_words;
  uint64_t data[kChunkWords];
};

__global__ void stage_records(uint64_t* __restrict__ dst, StageChunk c) {
  for (uint32_t i = threadIdx.x; i < c.n_words; i += blockDim.x) dst[c.offset_words + i] = c.data[i];
}

// Tile t covers events [t tile, min(n, (t + 1) tile)) of the single segment.
__global__ void stage_tiles(TileDesc* __restrict__ tiles, int n_tiles, int64_t n_events, int tile) {
  const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t < n_tiles) {
    TileDesc td;
    td.seg = 0;
    td.e0 = (int64_t)t * tile;
    const int64_t left = n_events - td.e0;
    td.count = (int32_t)(left < tile ? left : tile);
    tiles[t] = td;
  }
}

// The device address of `scratch` on the current device.
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_error(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return set_error(GPU_PHS_E_ARG, "generate_noalloc: device index beyond 63");
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)");
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return GPU_PHS_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

}  // namespace gphs

using namespace gphs;

extern "C" {

int gpu_phs_generate_noalloc(int process, int channel, int64_t n_events, const double* x, double* p,
                             double* factors, double* volumes, uint8_t* oks, void* stream_v) {
  std::lock_guard<std::mutex> lk(api_mutex());
  ProcessObj* po = find_process(process);
  if (!po) return set_error(GPU_PHS_E_HANDLE, "unknown process");
  const CompiledProcess& cp = po->cp;
  if (channel < 0 || channel >= (int)cp.channels.size()) return set_error(GPU_PHS_E_ARG, "channel out of range");
  if (n_events < 0) return set_error(GPU_PHS_E_ARG, "negative event count");
  if (n_events == 0) return ok_status();
  if (!x || !p || !factors || !volumes || !oks) return set_error(GPU_PHS_E_ARG, "buffers must not be NULL");
  int rc = require_device();
  if (rc) return rc;

  ChannelDesc cd = cp.channels[channel];
  const int n_nodes = cd.n_nodes;
  const NodeCmd* prog = &cp.nodes[cd.node0];
  cd.node0 = 0;
  cd.param0 = 0;
  const bool small = cd.shape >= 0 && kShapes[cd.shape].n_out <= 4;
  LaunchConfig lc;
  lc.block = small ? kBlockSmall : kBlockLarge;
  lc.ept = kEpt;
  const int tile = lc.block * lc.ept;
  const int64_t n_tiles = (n_events + tile - 1) / tile;
  if (n_tiles > kMaxTiles) {
    char msg[160];
    snprintf(msg, sizeof msg, "generate_noalloc: at most %lld events of this channel per call (got %lld)",
             (long long)(kMaxTiles * tile), (long long)n_events);
    return set_error(GPU_PHS_E_ARG, msg);
  }
  SegDesc sd{};
  sd.channel = 0;
  sd.n_events = n_events;
  sd.x = x; sd.p = p; sd.factor = factors; sd.volume = volumes; sd.ok = oks;
  sd.align16 = ((reinterpret_cast<uintptr_t>(x) % 16) == 0 ? 1 : 0) | ((reinterpret_cast<uintptr_t>(p) % 16) == 0 ? 2 : 0);
  const size_t o_seg = 0, o_chan = align256(sizeof(SegDesc)), o_nodes = o_chan + align256(sizeof(ChannelDesc));
  const size_t rec_bytes = o_nodes + sizeof(NodeCmd) * n_nodes;
  if (rec_bytes > kTableBytes) return set_error(GPU_PHS_E_ARG, "generate_noalloc: channel program too large");
  void* base = nullptr;
  rc = scratch_address(&base);
  if (rc) return rc;
  hipStream_t stream = (hipStream_t)stream_v;

  // the records, word by word through the staging kernel's arguments
  StageChunk chunk;
  unsigned char rec[kTableBytes];
  std::memset(rec, 0, rec_bytes);
  std::memcpy(rec + o_seg, &sd, sizeof sd);
  std::memcpy(rec + o_chan, &cd, sizeof cd);
  std::memcpy(rec + o_nodes, prog, sizeof(NodeCmd) * n_nodes);
  const size_t words = (rec_bytes + 7) / 8;
  for (size_t w0 = 0; w0 < words; w0 += kChunkWords) {
    chunk.offset_words = (uint32_t)w0;
    chunk.n_words = (uint32_t)std::min<size_t>(kChunkWords, words - w0);
    std::memcpy(chunk.data, rec + 8 * w0, 8 * chunk.n_words);
    hipLaunchKernelGGL(stage_records, dim3(1), dim3(256), 0, stream, (uint64_t*)base, chunk);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_error(e, "generate_noalloc: staging launch");
  }
  TileDesc* tiles = (TileDesc*)((unsigned char*)base + kTableBytes);
  hipLaunchKernelGGL(stage_tiles, dim3((unsigned)((n_tiles + 255) / 256)), dim3(256), 0, stream, tiles, (int)n_tiles,
                     n_events, tile);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_error(e, "generate_noalloc: tile launch");

  DevicePlan dp;
  dp.tiles = tiles;
  dp.segs = (const SegDesc*)((unsigned char*)base + o_seg);
  dp.chans = (const ChannelDesc*)((unsigned char*)base + o_chan);
  dp.nodes = (const NodeCmd*)((unsigned char*)base + o_nodes);
  dp.n_tiles = (int)n_tiles;
  dp.n_banks = cd.n_banks;
  dp.nx_max = cd.n_x;
  dp.nout_max = cd.n_out;
  dp.root_banks_only = 1;
  for (int i = 0; i < n_nodes; i++) {
    if (i > 0 && prog[i].push >= 0) dp.root_banks_only = 0;
    if (prog[i].friend_sign != 0) { dp.has_friends = 1; if (i == 0) dp.root_banks_only = 0; }
  }
  if (forward_lds_bytes(dp, lc) > forward_lds_limit()) return set_error(GPU_PHS_E_ARG, "generate_noalloc: LDS");
  e = cd.shape >= 0 ? launch_forward_shape(dp, cd.shape, 0, (int)n_tiles, lc, stream)
                    : launch_forward(dp, 0, (int)n_tiles, lc, stream);
  if (e != hipSuccess) return hip_error(e, "generate_noalloc: forward launch");
  return ok_status();
}

}  // extern "C"
