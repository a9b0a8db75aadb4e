// The device's streaming copy bandwidth: a practical ceiling for the HBM roofline of the
// generator's kernels, next to the published peak.  No counterpart in the reference.
//
// One thread copies one 16-byte word (one workgroup per 4 KB, so the launch has hundreds of
// thousands of short workgroups and the memory system sees as many requests as it can take),
// with streaming stores like the generator's output stores.  Reads and writes count alike.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "phs_device.h"
#include "phs_state.h"

using namespace gphs;

namespace {

__global__ __launch_bounds__(256) void stream_copy(const dvec2* __restrict__ in, gdouble2* __restrict__ out,
                                                   int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(in[i], out + i);
}

}  // namespace

extern "C" int gpu_phs_copy_bandwidth(int64_t bytes, int reps, void* stream, double* gb_s) {
  if (bytes < 4096 || bytes > ((int64_t)1 << 36) || reps < 1 || !gb_s)
    return set_error(GPU_PHS_E_ARG, "copy_bandwidth: need 4 KiB <= bytes <= 64 GiB, reps >= 1, gb_s");
  const int64_t n = bytes / 16;
  const int64_t blocks = (n + 255) / 256;
  if (blocks > 0x7fffffff) return set_error(GPU_PHS_E_ARG, "copy_bandwidth: too many workgroups");
  hipStream_t st = (hipStream_t)stream;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, n * 16);
  if (e == hipSuccess) e = hipMalloc(&b, n * 16);
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, n * 16, st);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  std::vector<float> t;
  for (int r = 0; e == hipSuccess && r <= reps; r++) {   // the first copy warms up
    e = hipEventRecord(e0, st);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(stream_copy, dim3((unsigned)blocks), dim3(256), 0, st, (const dvec2*)a, (gdouble2*)b, n);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && r > 0) t.push_back(ms);
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (e != hipSuccess) return hip_error(e, "copy_bandwidth");
  std::sort(t.begin(), t.end());
  const double ms = t[t.size() / 2];
  *gb_s = 2.0 * (double)n * 16 / (ms * 1e-3) / 1e9;
  return ok_status();
}
