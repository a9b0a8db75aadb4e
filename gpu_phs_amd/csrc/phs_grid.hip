// Randoms on the device, drawn from each channel's adaptive grid, and the grid's update from
// the accepted events (the reference's rng.cpp:18-93 on the host: std::default_random_engine
// through a std::piecewise_constant_distribution per (channel, dimension), 10 equal bins whose
// weights are the accepted events' histogram).
//
// The uniform numbers come from Philox4x32-10 (J. K. Salmon, M. A. Moraes, R. O. Dror,
// D. E. Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC 2011), a counter-based
// generator: the numbers of event e of segment s are a function of (seed, s, e, block), so
// every lane draws its own without state and a launch is reproducible whatever its shape.
// Two 32-bit outputs give one double u = (a 2^32 + b) >> 11 times 2^-53 in [0, 1).  u is
// mapped through the inverse distribution function of the piecewise-constant density of the
// bins: x = b_k + (u - F_k) / rho_k for the first bin k whose F_{k+1} > u.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "phs_device.h"
#include "phs_kernels.h"

namespace gphs {

namespace {

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    // one 32 x 32 -> 64-bit product gives both halves
    const uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += W0; k1 += W1;
  }
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  return (double)((((uint64_t)a << 32) | b) >> 11) * 0x1.0p-53;
}

// cdf: [channel][dim][kGridBins + 1] cumulative bin probabilities of the plan's channels.  One
// wave per workgroup: it copies its tile's channel table into LDS, then draws the tile's events
// 64 at a time.  The bin is the number of interior bounds F[1..9] at or below u, counted
// without branches.  Each lane puts its event's randoms into its column of an LDS table
// x[k][lane]; the wave then stores the 64 rows as one contiguous run (16-byte words where the
// row length is even and the rows are 16-byte aligned), so the writes leave as whole lines.
constexpr int kGenLanes = 64;

__global__ __launch_bounds__(kGenLanes) void grid_generate_kernel(const TileDesc* __restrict__ tiles, int n_tiles,
                                                                  const SegDesc* __restrict__ segs,
                                                                  const ChannelDesc* __restrict__ chans,
                                                                  const double* __restrict__ cdf, int n_dim,
                                                                  uint64_t seed) {
  extern __shared__ double cdf_l[];   // n_dim * (kGridBins + 1), then n_dim rows of kGenLanes + 1
  constexpr int NB1 = kGridBins + 1;
  constexpr int S = kGenLanes + 1;
  double* xs = cdf_l + n_dim * NB1;
  const int lane = (int)threadIdx.x;
  for (int t = (int)blockIdx.x; t < n_tiles; t += (int)gridDim.x) {
    const TileDesc td = tiles[t];
    const SegDesc sd = segs[td.seg];
    const int n_x = chans[sd.channel].n_x;
    const double* F0 = cdf + (size_t)sd.channel * n_dim * NB1;
    __syncthreads();   // the previous tile's tables have been read
    for (int i = lane; i < n_x * NB1; i += kGenLanes) cdf_l[i] = F0[i];
    __syncthreads();
    const bool vec = (sd.align16 & 1) && (n_x % 2 == 0);
    for (int base = 0; base < td.count; base += kGenLanes) {
      const int cnt = min(kGenLanes, td.count - base);
      const int64_t e = td.e0 + base + lane;
      if (lane < cnt) {
        for (int j = 0; 2 * j < n_x; j++) {
          uint32_t c[4] = {(uint32_t)e, (uint32_t)((uint64_t)e >> 32), (uint32_t)td.seg, (uint32_t)j};
          philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
          for (int h = 0; h < 2; h++) {
            const int k = 2 * j + h;
            if (k >= n_x) break;
            const double u = u53(c[2 * h], c[2 * h + 1]);
            const double* F = cdf_l + k * NB1;
            int b = 0;
#pragma unroll
            for (int q = 1; q < kGridBins; q++) b += F[q] <= u ? 1 : 0;
            const double rho = (F[b + 1] - F[b]) * kGridBins;   // density of bin b (width 1 / kGridBins)
            xs[k * S + lane] = (double)b / kGridBins + (u - F[b]) / rho;
          }
        }
      }
      __syncthreads();
      // rows [e0 + base, e0 + base + cnt) of x are cnt * n_x consecutive doubles
      const int n = cnt * n_x;
      double* dst = const_cast<double*>(sd.x) + (td.e0 + base) * (int64_t)n_x;
      if (vec) {
        gdouble2* d2 = (gdouble2*)dst;
        for (int w = lane; 2 * w < n; w += kGenLanes) {
          const int i = 2 * w, ev = i / n_x, k = i - ev * n_x;   // k even, so k + 1 < n_x
          dvec2 v;
          v.x = xs[k * S + ev];
          v.y = xs[(k + 1) * S + ev];
          d2[w] = v;
        }
      } else {
        gdouble* d1 = as_global(dst);
        for (int i = lane; i < n; i += kGenLanes) {
          const int ev = i / n_x;
          d1[i] = xs[(i - ev * n_x) * S + ev];
        }
      }
      __syncthreads();   // the table is read before the next events overwrite it
    }
  }
}

// Histogram of the accepted events' randoms per (channel, dim, bin) and the accepted count per
// channel: a workgroup histogram in LDS (its tile is one channel), then one global add per bin.
__global__ __launch_bounds__(256) void grid_histogram_kernel(const TileDesc* __restrict__ tiles, int n_tiles,
                                                             const SegDesc* __restrict__ segs,
                                                             const ChannelDesc* __restrict__ chans, int n_dim,
                                                             unsigned long long* __restrict__ counts,
                                                             unsigned long long* __restrict__ n_ok) {
  extern __shared__ unsigned int hist[];   // n_dim * kGridBins bins, then the accepted count
  const int nb = n_dim * kGridBins;
  for (int t = (int)blockIdx.x; t < n_tiles; t += (int)gridDim.x) {
    const TileDesc td = tiles[t];
    const SegDesc sd = segs[td.seg];
    const int n_x = chans[sd.channel].n_x;
    for (int i = (int)threadIdx.x; i <= nb; i += (int)blockDim.x) hist[i] = 0;
    __syncthreads();
    for (int local = (int)threadIdx.x; local < td.count; local += (int)blockDim.x) {
      const int64_t e = td.e0 + local;
      if (!as_global(sd.ok)[e]) continue;
      const gdouble* xr = as_global(sd.x) + e * n_x;
      for (int k = 0; k < n_x; k++) {
        int b = (int)(xr[k] * kGridBins);   // rng.cpp:52
        b = b < 0 ? 0 : b > kGridBins - 1 ? kGridBins - 1 : b;
        atomicAdd(&hist[k * kGridBins + b], 1u);
      }
      atomicAdd(&hist[nb], 1u);
    }
    __syncthreads();
    for (int i = (int)threadIdx.x; i <= nb; i += (int)blockDim.x) {
      const unsigned int v = hist[i];
      if (v == 0) continue;
      if (i < nb) atomicAdd(&counts[(size_t)sd.channel * nb + i], (unsigned long long)v);
      else atomicAdd(&n_ok[sd.channel], (unsigned long long)v);
    }
    __syncthreads();
  }
}

__global__ void fill_u64_kernel(unsigned long long* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

}  // namespace

hipError_t launch_grid_generate(const DevicePlan& dp, const double* cdf, int n_dim, uint64_t seed, hipStream_t stream) {
  if (dp.n_tiles <= 0) return hipSuccess;
  const size_t shmem = ((size_t)n_dim * (kGridBins + 1) + (size_t)n_dim * (kGenLanes + 1)) * sizeof(double);
  hipLaunchKernelGGL(grid_generate_kernel, dim3(dp.n_tiles), dim3(kGenLanes), shmem, stream, dp.tiles, dp.n_tiles,
                     dp.segs, dp.chans, cdf, n_dim, seed);
  return hipGetLastError();
}

hipError_t launch_grid_histogram(const DevicePlan& dp, int n_chans, int n_dim, unsigned long long* counts,
                                 hipStream_t stream) {
  unsigned long long* n_ok = counts + (size_t)n_chans * n_dim * kGridBins;
  const size_t n = (size_t)n_chans * n_dim * kGridBins + n_chans;
  hipLaunchKernelGGL(fill_u64_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 1024)), dim3(256), 0, stream,
                     counts, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || dp.n_tiles <= 0) return e;
  const size_t shmem = ((size_t)n_dim * kGridBins + 1) * sizeof(unsigned int);
  hipLaunchKernelGGL(grid_histogram_kernel, dim3(dp.n_tiles), dim3(256), shmem, stream, dp.tiles, dp.n_tiles,
                     dp.segs, dp.chans, n_dim, counts, n_ok);
  return hipGetLastError();
}

}  // namespace gphs
