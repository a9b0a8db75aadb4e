"""Process descriptions in the reference's tree format.

A process is a set of channels ("trees") over bincode-indexed branches, exactly the
arrays the reference's Whizard interface takes (whizard_gpu_phs.cpp:57-87):

* ``d1``, ``d2``, ``has_children``: ``[n_channels][n_prt]`` int32, daughters as bincodes
  (0 = none), ``n_prt = 2**(n_in + n_out) - 1``, index = bincode - 1;
* ``map_ids``, ``map_masses``, ``map_widths``: ``[n_channels][n_prt_out]`` with
  ``n_prt_out = 2**n_out - 1``, index = bincode - 1 (map ids of phs.h:8-17);
* ``flv_masses``, ``flv_widths``: ``[n_in + n_out]``, incoming first.

Final particle ``j`` (0-based) has bincode ``2**j``; the root branch has bincode
``2**n_out - 1``.  ``tree_channel`` builds one channel from a nested-tuple tree.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

MAP_NO = 0
MAP_SCHANNEL = 1
MAP_TCHANNEL = 2
MAP_UCHANNEL = 3
MAP_RADIATION = 4
MAP_COLLINEAR = 5
MAP_INFRARED = 6
MAP_STEP_E = 11
MAP_STEP_H = 12
MAP_ONSHELL = 99

ALL_MAPS = (MAP_NO, MAP_SCHANNEL, MAP_TCHANNEL, MAP_UCHANNEL, MAP_RADIATION, MAP_COLLINEAR,
            MAP_INFRARED, MAP_STEP_E, MAP_STEP_H)


@dataclass
class Process:
    n_in: int
    n_out: int
    d1: np.ndarray
    d2: np.ndarray
    has_children: np.ndarray
    map_ids: np.ndarray
    map_masses: np.ndarray
    map_widths: np.ndarray
    flv_masses: np.ndarray
    flv_widths: np.ndarray
    sqrts: float
    friends: np.ndarray | None = None
    name: str = ""
    labels: list = field(default_factory=list)

    @property
    def n_channels(self) -> int:
        return int(self.d1.shape[0])

    @property
    def n_prt(self) -> int:
        return 2 ** (self.n_in + self.n_out) - 1

    @property
    def n_prt_out(self) -> int:
        return 2 ** self.n_out - 1

    @property
    def n_x(self) -> int:
        # (n_out - 2) msq randoms + 2 (n_out - 1) angle randoms (phs_cpu.cpp:180-209, 301-311)
        return 3 * self.n_out - 4

    def friends_or_zero(self) -> np.ndarray:
        if self.friends is None:
            return np.zeros_like(self.d1)
        return self.friends

    def subset(self, channels) -> "Process":
        idx = np.asarray(channels, dtype=np.int64)
        return Process(self.n_in, self.n_out, self.d1[idx].copy(), self.d2[idx].copy(),
                       self.has_children[idx].copy(), self.map_ids[idx].copy(),
                       self.map_masses[idx].copy(), self.map_widths[idx].copy(),
                       self.flv_masses.copy(), self.flv_widths.copy(), self.sqrts,
                       None if self.friends is None else self.friends[idx].copy(), self.name,
                       [self.labels[i] for i in idx] if self.labels else [])

    def ctypes_args(self):
        """Contiguous arrays in the order the C entry points take them."""
        c = np.ascontiguousarray
        return (c(self.d1, np.int32), c(self.d2, np.int32), c(self.has_children, np.int32),
                c(self.map_ids, np.int32), c(self.map_masses, np.float64),
                c(self.map_widths, np.float64), c(self.flv_masses, np.float64))


def _bincode(node) -> int:
    if isinstance(node, int):
        return 1 << node
    return _bincode(node[0]) | _bincode(node[1])


def tree_channel(n_in: int, n_out: int, tree, maps=None):
    """One channel from a nested tuple tree of final-particle indices.

    ``tree = ((0, 1), (2, 3))`` is root -> (k1 = {0,1}, k2 = {2,3}).  ``maps`` maps a
    node (tuple, or "root") to ``(map_id, mass, width)``; unlisted internal nodes get
    MAP_NO.  Returns (d1, d2, hc, map_ids, map_m, map_w) rows.
    """
    n_prt = 2 ** (n_in + n_out) - 1
    n_prt_out = 2 ** n_out - 1
    d1 = np.zeros(n_prt, np.int32)
    d2 = np.zeros(n_prt, np.int32)
    hc = np.zeros(n_prt, np.int32)
    mid = np.zeros(n_prt_out, np.int32)
    mm = np.zeros(n_prt_out, np.float64)
    mw = np.zeros(n_prt_out, np.float64)
    maps = maps or {}
    leaves = []

    def walk(node, is_root):
        if isinstance(node, int):
            leaves.append(node)
            return
        b = _bincode(node)
        d1[b - 1] = _bincode(node[0])
        d2[b - 1] = _bincode(node[1])
        hc[b - 1] = 1
        key = "root" if is_root else node
        spec = maps.get(key, maps.get(node, (MAP_NO, 0.0, 0.0)))
        mid[b - 1], mm[b - 1], mw[b - 1] = spec
        walk(node[0], False)
        walk(node[1], False)

    walk(tree, True)
    if sorted(leaves) != list(range(n_out)) or _bincode(tree) != n_prt_out:
        raise ValueError("tree must contain each final particle exactly once")
    return d1, d2, hc, mid, mm, mw


def tree_friends(n_in: int, n_out: int, tree, friends=None) -> np.ndarray:
    """The friends row of one channel: ``friends`` maps a node (tuple, or "root") to the
    incoming particle (1 or 2) whose beam gives that branch's polar axis; the row holds the
    incoming particle's bincode (``2**n_out`` or ``2**(n_out + 1)``) at the branch's index,
    0 elsewhere (whizard_gpu_phs.cpp:79-87, read by phs.cu:276-284)."""
    row = np.zeros(2 ** (n_in + n_out) - 1, np.int32)
    for node, which in (friends or {}).items():
        if which not in (1, 2):
            raise ValueError("a friend is incoming particle 1 or 2")
        b = _bincode(tree if node == "root" else node)
        row[b - 1] = 1 << (n_out + which - 1)
    return row


def make_process(n_in, n_out, trees, maps_list, flv_masses, sqrts, flv_widths=None, name="",
                 friends_list=None):
    rows = [tree_channel(n_in, n_out, t, m) for t, m in zip(trees, maps_list)]
    cols = list(zip(*rows))
    flv_masses = np.asarray(flv_masses, np.float64)
    if flv_widths is None:
        flv_widths = np.zeros_like(flv_masses)
    friends = None
    if friends_list is not None:
        friends = np.stack([tree_friends(n_in, n_out, t, f) for t, f in zip(trees, friends_list)])
    return Process(n_in, n_out, np.stack(cols[0]), np.stack(cols[1]), np.stack(cols[2]),
                   np.stack(cols[3]), np.stack(cols[4]), np.stack(cols[5]), flv_masses,
                   np.asarray(flv_widths, np.float64), float(sqrts), friends, name,
                   [repr(t) for t in trees])


def random_tree(rng: np.random.Generator, leaves):
    """Uniformly shuffled leaves, random split points: an ordered binary tree."""
    leaves = list(leaves)
    if len(leaves) == 1:
        return leaves[0]
    k = int(rng.integers(1, len(leaves)))
    return (random_tree(rng, leaves[:k]), random_tree(rng, leaves[k:]))


def internal_nodes(tree):
    out = []

    def walk(node):
        if isinstance(node, int):
            return
        out.append(node)
        walk(node[0])
        walk(node[1])

    walk(tree)
    return out


def tree_depth(tree) -> int:
    if isinstance(tree, int):
        return 0
    return 1 + max(tree_depth(tree[0]), tree_depth(tree[1]))
