"""Trees with friends against Whizard's own events (CPU, oracle).

``tests/golden/bwbw_friends.ref`` holds the 149 events of samples/bwbw.ref whose trees have
``hasfriends: 1``.  The .ref format carries only that flag, not which branch has which friend,
so the assignment below was found by trying, for every tree, each of the 27 ways of giving its
three internal branches no friend, incoming particle 1 or incoming particle 2: Whizard's
all-channel factors single out the friend of the inner t/u-channel branch (the only branch of
these trees whose cos-theta map is peaked and whose polar axis is not the beam axis already),
and with it the forward pass reproduces Whizard's momenta from its randoms as closely as for
trees without friends.  The two other branches change neither: the root's axis is the beam
axis (friend +1 is the plain composition) and the remaining branch has a flat cos-theta map
and no friend.  Trees without friends, or with the other beam, miss by far more.
"""
import os

import numpy as np
import pytest

import oracle
from refio import read_ref

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

# 0-based channel -> (bincode of the branch, incoming particle 1 or 2)
FRIENDS = {43: (10, 1), 44: (7, 1), 45: (10, 2), 46: (13, 2), 47: (10, 1), 48: (7, 1), 49: (10, 2),
           50: (13, 2), 51: (10, 1), 52: (7, 1), 53: (10, 2), 54: (13, 2), 55: (7, 1), 56: (13, 2),
           57: (7, 1), 58: (13, 2)}


def _load(assign=FRIENDS):
    proc, ev = read_ref(os.path.join(GOLDEN, "bwbw_friends.ref"))
    proc.friends = np.zeros_like(proc.d1)
    for c, (b, which) in assign.items():
        proc.friends[c, b - 1] = 1 << (proc.n_out + which - 1)
    return proc, ev


def test_fixture_covers_every_tree_with_friends():
    proc, ev = _load()
    assert sorted(np.nonzero(ev["hasfriends"])[0]) == sorted(FRIENDS)
    assert set(ev["channel"]) == set(FRIENDS)
    assert len(ev["channel"]) == 149


def _momentum_dev(proc, ev):
    r = oracle.forward(proc, ev["x"], ev["channel"])
    p_w = ev["p"][:, proc.n_in:, :]
    return r, np.abs(r["p"] - p_w).max(axis=(1, 2)) / proc.sqrts


def test_friend_momenta_match_whizard():
    proc, ev = _load()
    r, dev = _momentum_dev(proc, ev)
    assert r["ok"].all()
    assert dev.max() < 1e-6, dev.max()
    assert np.allclose(r["volume"], ev["volume"], rtol=1e-12)


@pytest.mark.parametrize("variant", ["no friends", "other beam"])
def test_friend_momenta_need_the_right_beam(variant):
    if variant == "no friends":
        assign = {}
    else:
        assign = {c: (b, 3 - w) for c, (b, w) in FRIENDS.items()}
    proc, ev = _load(assign)
    _, dev = _momentum_dev(proc, ev)
    assert np.median(dev) > 1e-3


def test_friend_factors_match_whizard_all_channels():
    """Factors of every channel of every event (the inverse mapping), against Whizard's
    all_factors printed with 7 decimals."""
    proc, ev = _load()
    f, _ = oracle.inverse(proc, ev["p"][:, proc.n_in:, :])
    fa = np.array(ev["factors"])
    assert np.abs(f - fa).max() < 6e-8
