// The structure of a channel's program as the kernels walk it: read from the program records
// (SID < 0: generic kernels), or taken from the compile-time shape table (SID >= 0: kernels
// specialised for one tree shape, phs_shapes.h), where after the branch loops are unrolled every
// flag, random index and link below is a constant.
#pragma once
#include <hip/hip_runtime.h>

#include "phs_shapes.h"
#include "phs_types.h"

namespace gphs {

template <int SID>
struct Walk {
  static constexpr int kS = SID < 0 ? 0 : SID;
  static __host__ __device__ __forceinline__ int n_nodes(const ChannelDesc& cd) {
    if constexpr (SID >= 0) return kShapes[kS].n_nodes;
    else return cd.n_nodes;
  }
  // daughter c of branch i is a final-state particle
  static __host__ __device__ __forceinline__ bool leaf(const NodeCmd& nc, int i, int c) {
    if constexpr (SID >= 0) return c == 0 ? kShapes[kS].node[i].leaf0 : kShapes[kS].node[i].leaf1;
    else return nc.k[c].leaf;
  }
  // daughter visited next / deferred (-1: none)
  static __host__ __device__ __forceinline__ int desc(const NodeCmd& nc, int i) {
    if constexpr (SID >= 0) return kShapes[kS].node[i].desc;
    else return nc.desc;
  }
  static __host__ __device__ __forceinline__ int push(const NodeCmd& nc, int i) {
    if constexpr (SID >= 0) return kShapes[kS].node[i].push;
    else return nc.push;
  }
  // msq random of internal daughter c; phi random of the branch (cos theta is the next one)
  static __host__ __device__ __forceinline__ int xi(const NodeCmd& nc, int i, int c) {
    if constexpr (SID >= 0) return c == 0 ? kShapes[kS].node[i].xi0 : kShapes[kS].node[i].xi1;
    else return nc.k[c].xi;
  }
  static __host__ __device__ __forceinline__ int xphi(const NodeCmd& nc, int i) {
    if constexpr (SID >= 0) return kShapes[kS].node[i].xphi;
    else return nc.x_phi;
  }
  // visit index of internal daughter c's branch; of the parent branch (-1: the root)
  static __host__ __device__ __forceinline__ int child(const NodeCmd& nc, int i, int c) {
    if constexpr (SID >= 0) return c == 0 ? kShapes[kS].node[i].node0 : kShapes[kS].node[i].node1;
    else return nc.k[c].node;
  }
  static __host__ __device__ __forceinline__ int parent(const NodeCmd* prog, int i) {
    if constexpr (SID >= 0) return kShapes[kS].node[i].parent;
    else return prog[i].parent;
  }
};

}  // namespace gphs
