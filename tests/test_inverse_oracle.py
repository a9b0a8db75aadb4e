"""The inverse mapping in the oracle (CPU): the randoms and factor every channel would give an
event (phs.cu:626-675, 924-1284).

Pinned by
* the round trip: inverting an event into the channel that generated it returns its randoms
  and its factor, for every map id, for friends, and for unequal beams;
* Whizard's own all-channel factors (``all_factors`` of samples/ttH.ref and samples/bwbw.ref,
  printed with 4 and 7 decimals), computed from Whizard's momenta, for every channel;
* planted mistakes: the inverse of the wrong channel, or with the leaves of a pair swapped,
  misses Whizard's factors.
"""
import os

import numpy as np
import pytest

import oracle
from gpu_phs_amd import workloads as W
from refio import read_ref

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _golden_proc(name):
    from test_gpu_forward import load_golden   # noqa: the loader only reads the fixture
    proc, x, ch, _ = load_golden(os.path.join(GOLDEN, f"golden_{name}.npz"))
    return proc


PROCS = {
    "c0": W.process_c0, "c1": W.process_c1, "c2": W.process_c2, "2to6": W.process_2to6,
    "friends4": lambda: W.process_friends(4), "friends6": lambda: W.process_friends(6),
    "edge": lambda: _golden_proc("edge"),
}


def _phi_columns(proc):
    return [proc.n_out - 2 + 2 * k for k in range(proc.n_out - 1)]


def round_trip(proc, n=3000, seed=1, E=None):
    rng = np.random.default_rng(seed)
    x = rng.random((n, proc.n_x))
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    E1 = E2 = None
    if E is not None:
        E1, E2 = np.full(n, E[0]), np.full(n, E[1])
    o = oracle.forward(proc, x, ch, E1=E1, E2=E2)
    f, xo = oracle.inverse(proc, o["p"], E1=E1, E2=E2)
    idx = np.nonzero(o["ok"] & (o["factor"] > 0))[0]
    xr = xo[ch[idx], idx]
    dx = np.abs(xr - x[idx])
    phi = _phi_columns(proc)
    dx[:, phi] = np.minimum(dx[:, phi], 1 - dx[:, phi])     # phi is periodic in x
    df = np.abs(f[idx, ch[idx]] / o["factor"][idx] - 1)
    return idx, dx, df


@pytest.mark.parametrize("name", sorted(PROCS))
def test_inverse_of_the_generating_channel_returns_the_randoms(name):
    proc = PROCS[name]()
    idx, dx, df = round_trip(proc)
    assert len(idx) > 100
    # most events to the last bits: the msq of a branch is formed from momenta of size
    # sqrt(s), so a light branch (massless trees, t-channel maps) carries an error of
    # eps s / msq into its random and factor, and cos theta near +-1 of peaked maps loses
    # digits in both directions
    assert np.quantile(dx.max(axis=1), 0.5) < 1e-13
    assert np.quantile(dx.max(axis=1), 0.99) < 1e-11
    assert np.quantile(df, 0.99) < 1e-10
    assert dx.max() < 1e-8 and df.max() < 1e-7


def test_inverse_round_trip_unequal_beams():
    proc = W.process_friends(4)
    e1 = 0.65 * proc.sqrts
    idx, dx, df = round_trip(proc, E=(e1, proc.sqrts ** 2 / (4 * e1)))
    assert np.quantile(dx.max(axis=1), 0.99) < 1e-11 and np.quantile(df, 0.99) < 1e-10
    assert dx.max() < 1e-8 and df.max() < 1e-7


@pytest.mark.parametrize("name,atol", [("ttH_head.ref", 5.01e-5), ("bwbw_head.ref", 5.01e-8)])
def test_inverse_factors_match_whizard(name, atol):
    """All channels without friends (their per-branch friend data is not in the file;
    test_whizard_friends.py covers them)."""
    proc, ev = read_ref(os.path.join(GOLDEN, name))
    f, _ = oracle.inverse(proc, ev["p"][:, proc.n_in:, :])
    fa = np.array(ev["factors"])
    cols = ev["hasfriends"] == 0
    assert cols.sum() >= 3
    assert np.abs(f[:, cols] - fa[:, cols]).max() < atol


def test_inverse_planted_mistakes_miss_whizard():
    proc, ev = read_ref(os.path.join(GOLDEN, "bwbw_head.ref"))
    p = ev["p"][:, proc.n_in:, :]
    fa = np.array(ev["factors"])
    cols = ev["hasfriends"] == 0
    # leaves 0 and 2 swapped: the trees pair the final particles differently
    f, _ = oracle.inverse(proc, p[:, [2, 1, 0, 3], :])
    assert np.abs(f[:, cols] - fa[:, cols]).max() > 1e-3
    # the channels' factors rotated by one column
    f, _ = oracle.inverse(proc, p)
    assert np.abs(np.roll(f, 1, axis=1)[:, cols] - fa[:, cols]).max() > 1e-3
