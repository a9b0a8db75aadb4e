// Shape-specialised forward kernels of shapes [50, 64) of the shape table (phs_shapes.h);
// one translation unit per range so that they compile in parallel.  Trees of up to four
// final-state particles have variants for workgroups of 64, 128 and 256 threads and every
// register budget; those of five and six, for the 64-thread workgroups their launches take,
// at the compiler's budget, looping or one-pass (one workgroup per tile of one sub-tile: 150-160
// instead of 199-205 VGPRs, three waves per SIMD instead of two); other workgroup sizes run
// the variant with a run-time size.
#include "phs_forward.h"

namespace gphs {
namespace fwd {

namespace {

template <int SID, int BLK>
FwdKernel pick_wpe(int wpe) {
  if (wpe == 3) return shape_kernel<SID, 3, BLK>();
  if (wpe == 4) return shape_kernel<SID, 4, BLK>();
  if (wpe == 5) return shape_kernel<SID, 5, BLK>();
  return shape_kernel<SID, 0, BLK>();
}

template <int SID>
FwdKernel pick_one(int wpe, int block, bool one) {
  if constexpr (kShapes[SID].n_out <= 4) {
    if (block == 64) return pick_wpe<SID, 64>(wpe);
    if (block == 128) return pick_wpe<SID, 128>(wpe);
    if (block == 256) return pick_wpe<SID, 256>(wpe);
    return pick_wpe<SID, 0>(wpe);
  } else {
    if (block == 64) return one ? shape_kernel<SID, 0, 64, true>() : shape_kernel<SID, 0, 64>();
    return shape_kernel<SID, 0, 0>();
  }
}

template <int... S>
FwdKernel pick_seq(int sid, int wpe, int block, bool one, std::integer_sequence<int, S...>) {
  FwdKernel k = nullptr;
  (void)((sid == 50 + S ? (k = pick_one<50 + S>(wpe, block, one), true) : false) || ...);
  return k;
}

}  // namespace

FwdKernel pick_shape_50_64(int sid, int wpe, int block, bool one) {
  static_assert(64 <= kNumShapes, "shape range");
  return pick_seq(sid, wpe, block, one, std::make_integer_sequence<int, 14>{});
}

}  // namespace fwd
}  // namespace gphs
