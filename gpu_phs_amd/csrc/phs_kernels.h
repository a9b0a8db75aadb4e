// Launch interface between the host library and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "phs_types.h"

namespace gphs {

// Device copies of one plan's tables.
struct DevicePlan {
  const TileDesc* tiles = nullptr;
  const SegDesc* segs = nullptr;
  const ChannelDesc* chans = nullptr;
  const NodeCmd* nodes = nullptr;
  int n_tiles = 0;
  int n_banks = 0;       // save-bank depth the deepest program needs (0..3)
};

struct LaunchConfig {
  int block = 256;       // threads per workgroup (multiple of 64, <= 256)
  int ept = 1;           // events per thread per tile
  int grid = 0;          // workgroups; 0 = one per tile, else a persistent grid that strides
};

// Forward pass over tiles [t_begin, t_end).
hipError_t launch_forward(const DevicePlan& dp, int t_begin, int t_end, const LaunchConfig& lc, hipStream_t stream);

}  // namespace gphs
