// Compile-time tree shapes: the kernel programs (phs_types.h NodeCmd sequences) of every
// ordered binary tree with 2 to 6 final-state particles, reduced to what does not depend on
// the maps, masses or particle labels -- which daughters are final-state particles, which one
// the walk descends into and which one it defers, and the random index of every msq and angle.
//
// A channel whose program has one of these structures runs a forward kernel specialised for it
// (phs_forward.hip): the branch loop is unrolled, the randoms sit in registers addressed by
// constants and the walk's choices are resolved by the compiler, so no instruction selects a
// random by a run-time index or a daughter by a run-time flag.  The host compares each
// channel's program with this table entry by entry (shape_of); a channel that matches none runs
// the generic kernel, which reads the same fields from its program.
//
// The table (phs_shape_table.h) is written by gen_shapes.py: the programs tree_compiler.cpp
// emits for the trees named in its comments (preorder: 1 = branch, 0 = final-state particle):
// random indices follow the reference's consumption order (msq randoms of the internal
// non-root branches in post-order, phs_cpu.cpp:180-209, then phi and cos theta per branch in
// pre-order, phs_cpu.cpp:301-311), and the walk descends into the daughter with fewer
// final-state particles, so no tree of up to seven particles defers more than one subtree at a
// time (banks <= 1).
#pragma once
#include <stdint.h>

#include "phs_types.h"

namespace gphs {

struct ShapeNode {
  int8_t leaf0, leaf1;  // daughter 0 / 1 is a final-state particle
  int8_t desc, push;    // daughter visited next / deferred (-1: none)
  int8_t xi0, xi1;      // msq random of an internal daughter (-1: final-state particle)
  int8_t xphi;          // phi random; cos theta is the next one
  int8_t parent;        // visit index of the parent branch (-1: the root)
  int8_t node0, node1;  // visit index of an internal daughter's branch (-1: final-state particle)
};

struct ShapeProg {
  int8_t n_out, n_nodes, n_x, banks, root_only;
  ShapeNode node[5];
};

}  // namespace gphs

#include "phs_shape_table.h"

namespace gphs {

static_assert(kShapeMaxNodes <= 5, "ShapeProg holds five branches");

// The shape of a channel's program, or -1 (trees with friends always run the generic kernel).
inline int shape_of(const ChannelDesc& cd, const NodeCmd* prog) {
  for (int s = 0; s < kNumShapes; s++) {
    const ShapeProg& sh = kShapes[s];
    if (sh.n_out != cd.n_out || sh.n_nodes != cd.n_nodes || sh.n_x != cd.n_x) continue;
    bool same = true;
    for (int i = 0; i < sh.n_nodes && same; i++) {
      const NodeCmd& nc = prog[i];
      const ShapeNode& sn = sh.node[i];
      same = nc.friend_sign == 0 && nc.k[0].leaf == sn.leaf0 && nc.k[1].leaf == sn.leaf1 && nc.desc == sn.desc &&
             nc.push == sn.push && nc.x_phi == sn.xphi && (sn.leaf0 || nc.k[0].xi == sn.xi0) &&
             (sn.leaf1 || nc.k[1].xi == sn.xi1) && nc.parent == sn.parent &&
             (sn.leaf0 || nc.k[0].node == sn.node0) && (sn.leaf1 || nc.k[1].node == sn.node1);
    }
    if (same) return s;
  }
  return -1;
}

}  // namespace gphs
