"""Write the golden fixtures under tests/golden/ (run in a container where the reference
tree is mounted; the fixtures are committed, the reference never travels).

1. ``<name>.npz`` golden vectors: seeded U(0,1) randoms through the reference's own
   phs_cpu.cpp (compiled by oracle/Makefile into oracle/_ref/libphs_ref.so), for the
   benchmark processes and for edge-case trees exercising every map id.  Each file holds the
   process arrays, the randoms, the channel of every event and the reference outputs
   (momenta of the final particles, factor, volume, ok, msq of every branch).
2. ``ttH_head.ref`` / ``bwbw_head.ref``: the reference's sample files samples/ttH.ref and
   samples/bwbw.ref cut to their header, trees and first events (Whizard's own momenta).
3. ``bwbw_friends.ref``: samples/bwbw.ref with only the events of its trees that have friends.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from gpu_phs_amd import workloads as W  # noqa: E402
from gpu_phs_amd.process import (ALL_MAPS, MAP_NO, MAP_SCHANNEL, MAP_STEP_E, MAP_STEP_H,  # noqa: E402
                                 MAP_TCHANNEL, make_process)

REF_SAMPLES = "/root/reference/samples"


def edge_process():
    """2->3 trees (massive and massless) with every supported map id on the internal branch
    and on the root, including parameters that drive events out of the kinematic region."""
    trees, maps = [], []
    for mid in ALL_MAPS:
        for root_map in (MAP_NO, MAP_TCHANNEL):
            trees.append((0, (1, 2)))
            maps.append({"root": (root_map, 30.0, 0.0), (1, 2): (mid, 91.2, 2.5)})
            trees.append(((0, 2), 1))
            maps.append({"root": (root_map, 30.0, 0.0), (0, 2): (mid, 350.0, 40.0)})
    return make_process(2, 3, trees, maps, [0.000511, 0.000511, 4.8, 0.0, 125.0], 800.0,
                        name="edge: 2->3 all maps")


def write_npz(name, proc, n, seed):
    rng = np.random.default_rng(seed)
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    x = rng.random((n, proc.n_x))
    # a few exact end points and near-threshold values
    x[:8] = np.array([0.0, 1.0, 0.5, 1e-12, 1 - 1e-12, 0.25, 0.75, 0.5 - 1e-15])[:, None] * np.ones(proc.n_x)
    r = oracle.ref_forward(proc, x, ch, want_msq=True)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), n_in=proc.n_in, n_out=proc.n_out,
                        d1=proc.d1, d2=proc.d2, has_children=proc.has_children, map_ids=proc.map_ids,
                        map_masses=proc.map_masses, map_widths=proc.map_widths,
                        flv_masses=proc.flv_masses, flv_widths=proc.flv_widths, sqrts=proc.sqrts,
                        x=x, channels=ch, p=r["p"], factor=r["factor"], volume=r["volume"],
                        ok=r["ok"], msq=r["msq"])
    print(f"{name}: {n} events, ok fraction {r['ok'].mean():.3f}")


def head_ref(src, dst, n_events):
    with open(src) as f:
        lines = f.readlines()
    hdr = {}
    for ln in lines[:6]:
        k, v = ln.split()
        hdr[k] = int(v)
    n_ch, n_out = hdr["n_channels:"], hdr["NOUT:"]
    n_tree_lines = 7 * n_ch + 2
    per_event = 5 + 2 + n_out
    keep = lines[:6 + n_tree_lines + per_event * n_events]
    with open(dst, "w") as f:
        f.writelines(keep)
    print(f"{dst}: {len(keep)} lines")


def friend_ref(src, dst):
    """The header, trees and flavours of ``src`` and those of its events whose channel has
    ``hasfriends: 1`` (Whizard's momenta and all-channel factors for trees with friends)."""
    with open(src) as f:
        lines = f.readlines()
    hdr = {}
    for ln in lines[:6]:
        k, v = ln.split()
        hdr[k] = int(v)
    n_ch, n_out = hdr["n_channels:"], hdr["NOUT:"]
    n_tree_lines = 7 * n_ch + 2
    friend = [int(lines[6 + 7 * c + 3].split()[1]) for c in range(n_ch)]
    per_event = 5 + 2 + n_out
    keep = lines[:6 + n_tree_lines]
    pos, n = 6 + n_tree_lines, 0
    while pos + per_event <= len(lines):
        blk = lines[pos:pos + per_event]
        pos += per_event
        if friend[int(blk[0].split()[1]) - 1]:
            keep += blk
            n += 1
    with open(dst, "w") as f:
        f.writelines(keep)
    print(f"{dst}: {n} events")


def main():
    if not oracle.ref_available():
        oracle.build()
    write_npz("golden_c0", W.process_c0(), 2000, 100)
    write_npz("golden_c1", W.process_c1(), 4000, 101)
    write_npz("golden_c2", W.process_c2(), 4000, 102)
    write_npz("golden_2to6", W.process_2to6(), 4000, 103)
    write_npz("golden_edge", edge_process(), 4000, 104)
    if os.path.isdir(REF_SAMPLES):
        head_ref(os.path.join(REF_SAMPLES, "ttH.ref"), os.path.join(HERE, "ttH_head.ref"), 300)
        head_ref(os.path.join(REF_SAMPLES, "bwbw.ref"), os.path.join(HERE, "bwbw_head.ref"), 150)
        friend_ref(os.path.join(REF_SAMPLES, "bwbw.ref"), os.path.join(HERE, "bwbw_friends.ref"))


if __name__ == "__main__":
    main()
