// Host work-list compiler (see tree_compiler.h).
#include "tree_compiler.h"

#include "phs_mapconst.h"
#include "phs_shapes.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <functional>

namespace gphs {

namespace {

using mc::DualMW;
using mc::map_constants_d;

const double kTwoPi2 = 39.47841760435743447534;   // mappings.h:6
const double kTwoPi5 = 9792.62991312900650440772; // mappings.h:7

bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }
int ilog2(int v) { int l = 0; while (v > 1) { v >>= 1; l++; } return l; }

bool msq_kind(int id, int32_t* kind) {
  switch (id) {
    case MAP_NO: *kind = MK_LINEAR; return true;
    case MAP_SCHANNEL: *kind = MK_BREIT_WIGNER; return true;
    case MAP_COLLINEAR: case MAP_INFRARED: case MAP_RADIATION: *kind = MK_EXP; return true;
    case MAP_TCHANNEL: case MAP_UCHANNEL: *kind = MK_TU; return true;
    case MAP_STEP_E: *kind = MK_STEP_E; return true;
    case MAP_STEP_H: *kind = MK_STEP_H; return true;
    default: return false;
  }
}

bool ct_kind(int id, int32_t* kind) {
  switch (id) {
    case MAP_NO: case MAP_SCHANNEL: case MAP_INFRARED: case MAP_RADIATION: case MAP_STEP_E: case MAP_STEP_H:
      *kind = CK_FLAT; return true;
    case MAP_COLLINEAR: case MAP_TCHANNEL: case MAP_UCHANNEL:
      *kind = CK_PEAKED; return true;
    default: return false;
  }
}

void map_constants(int id, double m, double w, double m_min, double m_tot, double sqrts, double a[3], double b[3]) {
  DualMW ad[3], bd[3];
  map_constants_d(id, m, w, m_min, m_tot, sqrts, ad, bd);
  for (int k = 0; k < 3; k++) { a[k] = ad[k].v; b[k] = bd[k].v; }
}

}  // namespace

double channel_volume(const ProcessSpec& sp, int c, double sqrts) {
  const int n_prt = sp.n_prt();
  std::function<double(int)> vol = [&](int b) -> double {
    int k1 = sp.d1[c * n_prt + b] - 1, k2 = sp.d2[c * n_prt + b] - 1;
    double v1 = sp.has_children[c * n_prt + k1] ? vol(k1) : 1;
    double v2 = sp.has_children[c * n_prt + k2] ? vol(k2) : 1;
    if (b == sp.root()) return v1 * v2 / (4 * kTwoPi5);
    return v1 * v2 * sqrts * sqrts / (4 * kTwoPi2);
  };
  return vol(sp.root());
}

std::string compile_process(const ProcessSpec& sp, CompiledProcess& out) {
  char msg[256];
  if (sp.n_in != 2) return "n_in must be 2 (the tree format assumes two incoming particles)";
  if (sp.n_out < 2 || sp.n_out > kMaxOut) return "n_out must be in [2, 12]";
  if (sp.n_channels < 1) return "n_channels must be >= 1";
  const int n_prt = sp.n_prt(), n_prt_out = sp.n_prt_out(), root = sp.root();
  const size_t nc = sp.n_channels;
  if (sp.d1.size() < nc * n_prt || sp.d2.size() < nc * n_prt || sp.has_children.size() < nc * n_prt ||
      sp.map_ids.size() < nc * n_prt_out || sp.map_masses.size() < nc * n_prt_out ||
      sp.map_widths.size() < nc * n_prt_out || (int)sp.flv_masses.size() < sp.n_in + sp.n_out)
    return "process arrays are too short";
  if (!(sp.E1 > 0) || !(sp.E2 > 0)) return "beam energies must be positive";

  out = CompiledProcess();
  out.spec = sp;
  const double sqrts = 2 * std::sqrt(sp.E1 * sp.E2);   // phs.cu:1316
  out.sqrts = sqrts;
  out.mass_sum.assign(nc * n_prt_out, 0.0);

  for (int c = 0; c < sp.n_channels; c++) {
    const int* D1 = &sp.d1[c * n_prt];
    const int* D2 = &sp.d2[c * n_prt];
    const int* HC = &sp.has_children[c * n_prt];
    // ---- validate the tree (bincode consistency) and count leaves per branch
    std::vector<int> leaves(n_prt, 0);
    int seen_mask = 0;
    std::function<std::string(int, int)> check = [&](int b, int depth) -> std::string {
      int code = b + 1;
      if (depth > sp.n_out) return "tree deeper than n_out";
      if (!HC[b]) {
        if (!is_pow2(code) || code >= (1 << sp.n_out)) {
          snprintf(msg, sizeof msg, "channel %d: branch %d has no children but is not a final particle", c + 1, code);
          return msg;
        }
        if (seen_mask & code) { snprintf(msg, sizeof msg, "channel %d: final particle %d appears twice", c + 1, code); return msg; }
        seen_mask |= code;
        leaves[b] = 1;
        return "";
      }
      int k1 = D1[b], k2 = D2[b];
      if (k1 <= 0 || k2 <= 0 || (k1 | k2) != code || (k1 & k2) != 0 || k1 > n_prt || k2 > n_prt) {
        snprintf(msg, sizeof msg, "channel %d: branch %d has inconsistent daughters %d, %d", c + 1, code, k1, k2);
        return msg;
      }
      std::string e = check(k1 - 1, depth + 1);
      if (!e.empty()) return e;
      e = check(k2 - 1, depth + 1);
      if (!e.empty()) return e;
      leaves[b] = leaves[k1 - 1] + leaves[k2 - 1];
      return "";
    };
    if (!HC[root]) { snprintf(msg, sizeof msg, "channel %d: root branch has no children", c + 1); return msg; }
    std::string e = check(root, 0);
    if (!e.empty()) return e;
    if (seen_mask != (1 << sp.n_out) - 1) { snprintf(msg, sizeof msg, "channel %d: tree misses final particles", c + 1); return msg; }

    // ---- mass sums (phs.cu:61-78)
    double* ms = &out.mass_sum[c * n_prt_out];
    std::function<void(int)> msum = [&](int b) {
      if (HC[b]) {
        int k1 = D1[b] - 1, k2 = D2[b] - 1;
        msum(k1); msum(k2);
        ms[b] = ms[k1] + ms[k2];
      } else {
        ms[b] = sp.flv_masses[ilog2(b + 1) + 2];
      }
    };
    msum(root);
    const double m_tot = ms[root];

    // ---- random indices: msq post-order (root excluded), angles pre-order
    std::vector<int> xi(n_prt, -1), xphi(n_prt, -1);
    int cnt = 0;
    std::function<void(int)> post = [&](int b) {
      int k1 = D1[b] - 1, k2 = D2[b] - 1;
      if (HC[k1]) post(k1);
      if (HC[k2]) post(k2);
      if (b != root) xi[b] = cnt++;
    };
    post(root);
    cnt = sp.n_out - 2;
    std::function<void(int)> pre = [&](int b) {
      xphi[b] = cnt; cnt += 2;
      int k1 = D1[b] - 1, k2 = D2[b] - 1;
      if (HC[k1]) pre(k1);
      if (HC[k2]) pre(k2);
    };
    pre(root);

    // ---- per-branch records
    auto mid = [&](int b) { return sp.map_ids[c * n_prt_out + b]; };
    // sum program of branch k's momentum (see ChildCmd::sum_code): a branch whose daughters
    // are both internal evaluates the one that needs the second accumulator first, then the
    // other one in the second accumulator, which it must not need; false if that fails
    std::function<bool(int)> needs_two = [&](int b) -> bool {
      return HC[b] && ((HC[D1[b] - 1] && HC[D2[b] - 1]) || needs_two(D1[b] - 1) || needs_two(D2[b] - 1));
    };
    std::function<bool(int, int, std::vector<int>&)> sum_ops = [&](int b, int level, std::vector<int>& ops) -> bool {
      if (!HC[b]) { ops.push_back(ilog2(b + 1)); return true; }
      int k1 = D1[b] - 1, k2 = D2[b] - 1;
      if (!HC[k1] || !HC[k2]) {
        // one or two final-state daughters: the internal one (if any) first, then add
        const int in = HC[k1] ? k1 : k2, lf = HC[k1] ? k2 : k1;
        if (!sum_ops(in, level, ops)) return false;
        ops.push_back(ilog2(lf + 1));
        return true;
      }
      if (level > 0 || (needs_two(k1) && needs_two(k2))) return false;
      const int first = needs_two(k2) ? k2 : k1, second = first == k1 ? k2 : k1;
      if (!sum_ops(first, 0, ops)) return false;
      ops.push_back(kSumPush);
      if (!sum_ops(second, 1, ops)) return false;
      ops.push_back(kSumMerge);
      return true;
    };
    auto fill_child = [&](int k, ChildCmd& cc) -> std::string {
      cc = ChildCmd();
      cc.leaves = (uint32_t)(k + 1);   // a bincode is the set of final-state particles below it
      std::vector<int> ops;
      if (sum_ops(k, 0, ops) && (int)ops.size() <= kSumMaxOps) {
        cc.sum_len = (int32_t)ops.size();
        for (int t = 0; t < cc.sum_len; t++) cc.sum_code |= (uint64_t)ops[t] << (4 * t);
      }
      if (!HC[k]) {
        int j = ilog2(k + 1);
        double mj = sp.flv_masses[sp.n_in + j];
        cc.leaf = 1; cc.out = j; cc.msq_leaf = mj * mj;   // phs_cpu.cpp:544
        cc.m_leaf = std::sqrt(cc.msq_leaf);
        cc.node = -1;
        return "";
      }
      cc.leaf = 0;
      cc.node = -1;
      if (!msq_kind(mid(k), &cc.kind)) {
        snprintf(msg, sizeof msg, "channel %d: branch %d has map id %d, which has no msq mapping", c + 1, k + 1, mid(k));
        return msg;
      }
      DualMW a[3], b[3];
      double m = sp.map_masses[c * n_prt_out + k], w = sp.map_widths[c * n_prt_out + k];
      map_constants_d(mid(k), m, w, ms[k], m_tot, sqrts, a, b);
      double m_max = sqrts - m_tot + ms[k];
      cc.xi = xi[k];
      cc.a0 = a[0].v; cc.a1 = a[1].v; cc.a2 = a[2].v;
      for (int q = 0; q < 3; q++) { cc.da_dm[q] = a[q].dm; cc.da_dw[q] = a[q].dw; }
      cc.m = m; cc.w = w;
      cc.m_min = ms[k];
      cc.msq_min = ms[k] * ms[k];
      cc.msq_max = m_max * m_max;
      cc.f_lin = (cc.msq_max - cc.msq_min) / (sqrts * sqrts);
      return "";
    };

    ChannelDesc cd{};
    cd.node0 = (int32_t)out.nodes.size();
    cd.n_nodes = sp.n_out - 1;
    cd.n_out = sp.n_out;
    cd.n_x = sp.n_x();
    cd.div_nx = (uint32_t)(((uint64_t)1 << 32) / (uint64_t)cd.n_x + 1);
    cd.div_p = (uint32_t)(((uint64_t)1 << 32) / (uint64_t)(4 * cd.n_out) + 1);
    cd.param0 = cd.node0;
    cd.sqrts = sqrts;
    cd.s = sqrts * sqrts;
    cd.volume = channel_volume(sp, c, sqrts);
    cd.bg_lab = (sp.E1 - sp.E2) / (2 * std::sqrt(sp.E1 * sp.E2));   // phs.cu:1289
    cd.g_lab = std::sqrt(1 + cd.bg_lab * cd.bg_lab);
    cd.m_tot = m_tot;

    // ---- visit order: descend into the smaller internal subtree, defer the other
    int depth = 0, max_depth = 0, n_tape = 0;
    std::function<std::string(int, int, int)> emit = [&](int b, int parent, int slot) -> std::string {
      NodeCmd nd{};
      int k[2] = {D1[b] - 1, D2[b] - 1};
      if (!ct_kind(mid(b), &nd.ct_kind)) {
        snprintf(msg, sizeof msg, "channel %d: branch %d has map id %d, which has no cos-theta mapping", c + 1, b + 1, mid(b));
        return msg;
      }
      // friend sign (phs.cu:276-284): the incoming particle with bincode 2^n_out gives -1,
      // the one with 2^(n_out + 1) gives +1, any other entry no friend
      const int fr = sp.friends.empty() ? 0 : sp.friends[c * n_prt + b];
      nd.friend_sign = fr == (1 << sp.n_out) ? -1 : fr == (1 << (sp.n_out + 1)) ? 1 : 0;
      if (nd.friend_sign) out.has_friends = true;
      DualMW a[3], bb[3];
      double m = sp.map_masses[c * n_prt_out + b], w = sp.map_widths[c * n_prt_out + b];
      map_constants_d(mid(b), m, w, ms[b], m_tot, sqrts, a, bb, nd.friend_sign ? 4 * sp.E1 * sp.E2 : 0);
      nd.x_phi = xphi[b];
      nd.map_id = mid(b);
      nd.parent = parent;
      nd.slot = slot;
      nd.m_max = (b == root) ? sqrts : sqrts - m_tot + ms[b];
      nd.b0 = bb[0].v; nd.b1 = bb[1].v;
      nd.db_dm[0] = bb[0].dm; nd.db_dm[1] = bb[1].dm;
      nd.m_min = ms[b];
      nd.m = m;
      for (int i = 0; i < 2; i++) {
        std::string err = fill_child(k[i], nd.k[i]);
        if (!err.empty()) return err;
      }
      bool in0 = HC[k[0]], in1 = HC[k[1]];
      nd.desc = -1; nd.push = -1;
      if (in0 && in1) {
        nd.desc = leaves[k[1]] < leaves[k[0]] ? 1 : 0;
        nd.push = 1 - nd.desc;
      } else if (in0) {
        nd.desc = 0;
      } else if (in1) {
        nd.desc = 1;
      }
      const int me = (int)out.nodes.size() - cd.node0;
      nd.tape_slot = (in0 || in1) ? n_tape++ : -1;
      out.nodes.push_back(nd);
      out.node_branch.push_back(b);
      if (nd.push >= 0) { depth++; max_depth = std::max(max_depth, depth); }
      if (nd.desc >= 0) {
        out.nodes[cd.node0 + me].k[nd.desc].node = (int)out.nodes.size() - cd.node0;
        std::string err = emit(k[nd.desc], me, nd.desc);
        if (!err.empty()) return err;
      }
      if (nd.push >= 0) {
        depth--;
        out.nodes[cd.node0 + me].k[nd.push].node = (int)out.nodes.size() - cd.node0;
        std::string err = emit(k[nd.push], me, nd.push);
        if (!err.empty()) return err;
      }
      return "";
    };
    e = emit(root, -1, 0);
    if (!e.empty()) return e;
    cd.n_banks = max_depth;
    cd.n_tape = n_tape;
    cd.shape = shape_of(cd, &out.nodes[cd.node0]);
    out.max_banks = std::max(out.max_banks, max_depth);
    if ((int)out.nodes.size() - cd.node0 != cd.n_nodes) return "internal error: node count";
    out.channels.push_back(cd);
  }
  return "";
}

}  // namespace gphs
