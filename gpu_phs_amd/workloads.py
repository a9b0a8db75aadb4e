"""Benchmark and test workloads: the five configurations this package is measured on.

  c0  2->2  e+e- -> t tbar at sqrt(s) = 500 GeV, final masses 173.0 GeV, one channel,
      root bincode 3 -> (1, 2) with MAP_NO; n_x = 2 (phi first, then cos theta); 1e6 events.
  c1  2->4 double resonance at sqrt(s) = 500 GeV, four massless finals, three channels:
      ((12)(34)) and ((14)(23)) with Breit-Wigner maps M = 80.4, Gamma = 2.1 on both
      resonances, and the t-channel tree (1,(2,(34))) with MAP_TCHANNEL (regulator 10 GeV)
      at the root, MAP_NO on (234), Breit-Wigner on (34); n_x = 8; 1e7 events.
  c2  2->4 massless multi-jet at sqrt(s) = 1 TeV: 120 channels from seeded random ordered
      binary trees, internal maps drawn from {INFRARED, COLLINEAR, TCHANNEL, UCHANNEL}
      with regulator mass 10 GeV, root map from {NO, TCHANNEL}; 1e7 events split evenly.
  c3  batched uneven mix: every channel of c0-c2 plus 16 2->6 trees (six 173 GeV
      finals at sqrt(s) = 3 TeV, seeded random trees including the depth-5 chain, with
      Breit-Wigner / step / t-channel maps); per-channel event counts drawn
      Dirichlet(0.5) from the seed; 5e7 events.
  c4  adjoint on the c1 trees, 1e6 events; loss = sum_events ok * factor * sum_j pT_j^2 / s.

Seeds: 0x9E3779B97F4A7C15 ^ config index.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .process import (MAP_COLLINEAR, MAP_INFRARED, MAP_NO, MAP_SCHANNEL, MAP_STEP_E, MAP_STEP_H,
                      MAP_TCHANNEL, MAP_UCHANNEL, Process, internal_nodes, make_process,
                      random_tree)

DESCRIPTIONS = {
    "c0": "2->2 e+e- -> t tbar at sqrt(s)=500 GeV, 1 channel, 2 randoms/event, 1e6 events",
    "c1": "2->4 double resonance at sqrt(s)=500 GeV (BW M=80.4 G=2.1 + t-channel tree), 3 channels, "
          "8 randoms/event, 1e7 events",
    "c2": "2->4 massless multi-jet at sqrt(s)=1 TeV, 120 random-tree channels, 8 randoms/event, 1e7 events",
    "c3": "batched uneven mix: channels of c0-c2 + 16 2->6 trees (6 x 173 GeV at 3 TeV), "
          "Dirichlet(0.5) per-channel counts, 5e7 events",
    "c4": "adjoint on the c1 trees, 1e6 events, loss = sum ok*factor*sum pT^2/s",
}

SEED_BASE = 0x9E3779B97F4A7C15
M_E = 0.000511


def seed(config_index: int) -> int:
    return (SEED_BASE ^ config_index) & 0xFFFFFFFFFFFFFFFF


def process_c0() -> Process:
    return make_process(2, 2, [(0, 1)], [{"root": (MAP_NO, 0.0, 0.0)}],
                        [M_E, M_E, 173.0, 173.0], 500.0, name="c0: e+e- -> t tbar")


def process_c1() -> Process:
    bw = (MAP_SCHANNEL, 80.4, 2.1)
    trees = [((0, 1), (2, 3)), ((0, 3), (1, 2)), (0, (1, (2, 3)))]
    maps = [
        {(0, 1): bw, (2, 3): bw},
        {(0, 3): bw, (1, 2): bw},
        {"root": (MAP_TCHANNEL, 10.0, 0.0), (1, (2, 3)): (MAP_NO, 0.0, 0.0), (2, 3): bw},
    ]
    return make_process(2, 4, trees, maps, [M_E, M_E, 0.0, 0.0, 0.0, 0.0], 500.0,
                        name="c1: 2->4 double resonance")


def process_c2(n_channels: int = 120) -> Process:
    rng = np.random.default_rng(seed(2))
    trees, maps = [], []
    internal = (MAP_INFRARED, MAP_COLLINEAR, MAP_TCHANNEL, MAP_UCHANNEL)
    for _ in range(n_channels):
        leaves = list(rng.permutation(4))
        t = random_tree(rng, [int(v) for v in leaves])
        m = {"root": ((MAP_NO, MAP_TCHANNEL)[int(rng.integers(2))], 10.0, 0.0)}
        for node in internal_nodes(t)[1:]:
            m[node] = (internal[int(rng.integers(4))], 10.0, 0.0)
        trees.append(t)
        maps.append(m)
    return make_process(2, 4, trees, maps, [M_E, M_E, 0.0, 0.0, 0.0, 0.0], 1000.0,
                        name="c2: 2->4 massless multi-jet")


def process_2to6(n_channels: int = 16) -> Process:
    rng = np.random.default_rng(seed(3) + 6)
    chain = (0, (1, (2, (3, (4, 5)))))
    trees = [chain, (((0, 1), (2, 3)), (4, 5))]
    while len(trees) < n_channels:
        trees.append(random_tree(rng, [int(v) for v in rng.permutation(6)]))
    choices = [(MAP_NO, 0.0, 0.0), (MAP_SCHANNEL, 420.0, 12.0), (MAP_STEP_E, 600.0, 30.0),
               (MAP_STEP_H, 700.0, 25.0), (MAP_TCHANNEL, 50.0, 0.0)]
    maps = []
    for t in trees:
        m = {"root": ((MAP_NO, MAP_TCHANNEL)[int(rng.integers(2))], 50.0, 0.0)}
        for node in internal_nodes(t)[1:]:
            m[node] = choices[int(rng.integers(len(choices)))]
        maps.append(m)
    return make_process(2, 6, trees, maps, [M_E, M_E] + [173.0] * 6, 3000.0,
                        name="2->6: six 173 GeV finals")


def process_large(n_out: int = 8, n_channels: int = 12) -> Process:
    """Test process with more final-state particles than the benchmarks: seeded random trees
    plus the chain and the balanced tree, 10 GeV finals at sqrt(s) = 4 TeV, internal maps
    drawn from {collinear, infrared, t-channel, Breit-Wigner, step E/H}.  Exercises save banks up to
    three levels deep, randoms in LDS (n_x > 16) and long daughter sum programs."""
    rng = np.random.default_rng(seed(7) + n_out)
    leaves = list(range(n_out))
    chain = leaves[-1]
    for j in reversed(leaves[:-1]):
        chain = (j, chain)
    def balanced(ls):
        return ls[0] if len(ls) == 1 else (balanced(ls[:len(ls) // 2]), balanced(ls[len(ls) // 2:]))
    trees = [chain, balanced(leaves)]
    while len(trees) < n_channels:
        trees.append(random_tree(rng, [int(v) for v in rng.permutation(n_out)]))
    # maps that favour light inner branches, so that a fair share of deep trees is kinematically
    # allowed (each branch must be heavier than its daughters)
    choices = [(MAP_COLLINEAR, 10.0, 0.0), (MAP_INFRARED, 10.0, 0.0), (MAP_TCHANNEL, 50.0, 0.0),
               (MAP_SCHANNEL, 120.0, 5.0), (MAP_STEP_E, 150.0, 10.0), (MAP_STEP_H, 150.0, 10.0)]
    maps = []
    for t in trees:
        m = {"root": ((MAP_NO, MAP_TCHANNEL)[int(rng.integers(2))], 50.0, 0.0)}
        for node in internal_nodes(t)[1:]:
            m[node] = choices[int(rng.integers(len(choices)))]
        maps.append(m)
    return make_process(2, n_out, trees, maps, [M_E, M_E] + [10.0] * n_out, 4000.0,
                        name=f"2->{n_out}: 10 GeV finals")


def ordered_trees(leaves):
    """Every ordered binary tree over the given leaves in that order."""
    if len(leaves) == 1:
        return [leaves[0]]
    out = []
    for k in range(1, len(leaves)):
        for a in ordered_trees(leaves[:k]):
            for b in ordered_trees(leaves[k:]):
                out.append((a, b))
    return out


def process_shapes(n_out: int, per_shape: int = 3) -> Process:
    """Every ordered tree shape with n_out final-state particles (2 to 6), each with
    ``per_shape`` seeded labellings of the particles and seeded maps from the whole set
    (Breit-Wigner, step, exponential, t/u-channel, linear), particles 1 and 4 massive (4.8 and
    1.5 GeV), the others massless, at sqrt(s) = 1 TeV: the trees of the shape-specialised
    forward kernels."""
    rng = np.random.default_rng(seed(9) + n_out)
    choices = [(MAP_NO, 0.0, 0.0), (MAP_SCHANNEL, 91.2, 2.5), (MAP_SCHANNEL, 300.0, 15.0),
               (MAP_COLLINEAR, 10.0, 0.0), (MAP_INFRARED, 10.0, 0.0), (MAP_TCHANNEL, 20.0, 0.0),
               (MAP_UCHANNEL, 20.0, 0.0), (MAP_STEP_E, 250.0, 20.0), (MAP_STEP_H, 250.0, 20.0)]
    trees, maps = [], []
    for shape in ordered_trees(list(range(n_out))):
        for _ in range(per_shape):
            perm = [int(v) for v in rng.permutation(n_out)]
            relabel = lambda t: perm[t] if isinstance(t, int) else (relabel(t[0]), relabel(t[1]))
            t = relabel(shape)
            m = {"root": ((MAP_NO, MAP_TCHANNEL, MAP_UCHANNEL)[int(rng.integers(3))], 30.0, 0.0)}
            for node in internal_nodes(t)[1:]:
                m[node] = choices[int(rng.integers(len(choices)))]
            trees.append(t)
            maps.append(m)
    flv = [M_E, M_E] + [4.8, 0.0, 0.0, 1.5, 0.0, 0.0][:n_out]
    return make_process(2, n_out, trees, maps, flv, 1000.0, name=f"2->{n_out}: every tree shape")


def process_friends(n_out: int = 4) -> Process:
    """Trees whose branches take their polar axis from an incoming beam (friends), at
    sqrt(s) = 1 TeV: friend signs -1 and +1, friends on the root and on inner branches, flat
    and peaked cos-theta maps on friend branches."""
    bw = (MAP_SCHANNEL, 80.4, 2.1)
    t_ch = (MAP_TCHANNEL, 10.0, 0.0)
    if n_out == 4:
        trees = [((0, 1), (2, 3)), (0, (1, (2, 3))), ((0, 3), (1, 2)), ((0, 1), (2, 3))]
        maps = [{(0, 1): bw, (2, 3): bw},
                {"root": t_ch, (1, (2, 3)): t_ch, (2, 3): bw},
                {(0, 3): bw, (1, 2): bw},
                {"root": t_ch, (0, 1): t_ch, (2, 3): (MAP_COLLINEAR, 10.0, 0.0)}]
        friends = [{"root": 1, (0, 1): 2},
                   {(1, (2, 3)): 1, (2, 3): 2},
                   {"root": 2},
                   {(0, 1): 1, (2, 3): 1}]
        return make_process(2, 4, trees, maps, [M_E, M_E, 0.0, 0.0, 5.0, 5.0], 1000.0,
                            name="2->4 with friends", friends_list=friends)
    if n_out == 6:
        chain = (0, (1, (2, (3, (4, 5)))))
        pairs = (((0, 1), (2, 3)), (4, 5))
        trees = [chain, pairs]
        maps = [{"root": t_ch, (1, (2, (3, (4, 5)))): t_ch, (2, (3, (4, 5))): (MAP_NO, 0.0, 0.0),
                 (3, (4, 5)): t_ch, (4, 5): (MAP_SCHANNEL, 420.0, 12.0)},
                {((0, 1), (2, 3)): (MAP_NO, 0.0, 0.0), (0, 1): (MAP_SCHANNEL, 420.0, 12.0),
                 (2, 3): t_ch, (4, 5): (MAP_SCHANNEL, 420.0, 12.0)}]
        friends = [{(1, (2, (3, (4, 5)))): 1, (2, (3, (4, 5))): 2, (3, (4, 5)): 1, (4, 5): 2},
                   {"root": 1, (0, 1): 2, (2, 3): 1}]
        return make_process(2, 6, trees, maps, [M_E, M_E] + [173.0] * 6, 3000.0,
                            name="2->6 with friends", friends_list=friends)
    raise ValueError("n_out must be 4 or 6")


@dataclass
class Segment:
    process: int      # index into Workload.processes
    channel: int
    n_events: int


@dataclass
class Workload:
    name: str
    processes: list
    segments: list
    seed: int

    @property
    def n_events(self) -> int:
        return sum(s.n_events for s in self.segments)

    def bytes_per_event(self, proc: Process) -> int:
        # x read + leaf momenta, factor, volume written + ok byte
        return 8 * proc.n_x + 32 * proc.n_out + 8 + 8 + 1

    def algorithmic_bytes(self) -> int:
        return sum(s.n_events * self.bytes_per_event(self.processes[s.process]) for s in self.segments)


def _even(n_events, n_channels):
    base = n_events // n_channels
    counts = [base] * n_channels
    for i in range(n_events - base * n_channels):
        counts[i] += 1
    return counts


def workload(name: str, scale: float = 1.0) -> Workload:
    """Workload by config name ("c0".."c4"); ``scale`` shrinks the event counts for tests."""
    if name == "c0":
        n = max(1, int(1_000_000 * scale))
        return Workload(name, [process_c0()], [Segment(0, 0, n)], seed(0))
    if name in ("c1", "c4"):
        n = max(3, int((10_000_000 if name == "c1" else 1_000_000) * scale))
        segs = [Segment(0, c, k) for c, k in enumerate(_even(n, 3))]
        return Workload(name, [process_c1()], segs, seed(1 if name == "c1" else 4))
    if name == "c2":
        n = max(120, int(10_000_000 * scale))
        segs = [Segment(0, c, k) for c, k in enumerate(_even(n, 120))]
        return Workload(name, [process_c2()], segs, seed(2))
    if name == "c3":
        procs = [process_c0(), process_c1(), process_c2(), process_2to6()]
        chans = [(p, c) for p, pr in enumerate(procs) for c in range(pr.n_channels)]
        rng = np.random.default_rng(seed(3))
        w = rng.dirichlet(np.full(len(chans), 0.5))
        n = int(50_000_000 * scale)
        counts = np.floor(w * n).astype(np.int64)
        counts[np.argmax(counts)] += n - int(counts.sum())
        segs = [Segment(p, c, int(k)) for (p, c), k in zip(chans, counts) if k > 0]
        return Workload(name, procs, segs, seed(3))
    raise ValueError(f"unknown workload {name!r}")
