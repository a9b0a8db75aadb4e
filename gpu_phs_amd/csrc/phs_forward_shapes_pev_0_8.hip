// Shape-specialised forward kernels for per-event beam energies, shapes [0, 8) of the shape
// table (the trees of 2-4 final-state particles): each event's sqrt(s), lab boost and
// sqrt(s)-dependent map constants formed in the kernel, as in the generic kernel's per-event
// variant, with the walk unrolled for the shape.  Workgroups of 64, 128 and 256 threads (or a
// run-time size) at the compiler's register budget or one for 3 waves per SIMD (asked for 4 or
// 5, the kernel takes 3), and one-pass variants (one workgroup per tile of one sub-tile) at the
// compiler's budget, 106-128 VGPRs.  Larger trees with per-event beams run the generic kernel.
#include "phs_forward.h"

namespace gphs {
namespace fwd {

namespace {

template <int SID, int WPE, int BLK, bool ONE = false>
FwdKernel pev_kernel() {
  constexpr const ShapeProg& sh = kShapes[SID];
  constexpr int bank = sh.banks == 0 ? BK_NONE : (sh.root_only ? BK_ROOT : BK_VGPR);
  return forward_kernel<bank, WPE, sh.n_x, false, true, SID, BLK, ONE>;
}

template <int SID, int BLK>
FwdKernel pick_wpe(int wpe) {
  return wpe >= 3 ? pev_kernel<SID, 3, BLK>() : pev_kernel<SID, 0, BLK>();
}

template <int SID>
FwdKernel pick_one(int wpe, int block, bool one) {
  static_assert(kShapes[SID].n_out <= 4, "per-event beam shape kernels cover trees of up to four particles");
  if (one) {
    if (block == 64) return pev_kernel<SID, 0, 64, true>();
    if (block == 128) return pev_kernel<SID, 0, 128, true>();
    if (block == 256) return pev_kernel<SID, 0, 256, true>();
  }
  if (block == 64) return pick_wpe<SID, 64>(wpe);
  if (block == 128) return pick_wpe<SID, 128>(wpe);
  if (block == 256) return pick_wpe<SID, 256>(wpe);
  return pick_wpe<SID, 0>(wpe);
}

template <int... S>
FwdKernel pick_seq(int sid, int wpe, int block, bool one, std::integer_sequence<int, S...>) {
  FwdKernel k = nullptr;
  (void)((sid == S ? (k = pick_one<S>(wpe, block, one), true) : false) || ...);
  return k;
}

}  // namespace

FwdKernel pick_shape_pev_0_8(int sid, int wpe, int block, bool one) {
  return pick_seq(sid, wpe, block, one, std::make_integer_sequence<int, kPevShapes>{});
}

}  // namespace fwd
}  // namespace gphs
