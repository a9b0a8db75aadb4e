"""GPU parity of the friend axis (branches whose polar axis is an incoming beam) against the
oracle, through the C ABI.  The oracle's friend construction is pinned by the invariants of
test_friend_oracle.py; the reference's own version is GPU-only with a non-orthogonal rotation,
so parity with the reference itself is unpinned."""
import numpy as np
import pytest

import oracle
from gpu_phs_amd import workloads as W
from parity import assert_parity, forward_check
from test_gpu_forward import run_gpu

pytestmark = pytest.mark.gpu


def _events(proc, n, seed):
    rng = np.random.default_rng(seed)
    return rng.random((n, proc.n_x)), rng.integers(0, proc.n_channels, n).astype(np.int32)


@pytest.mark.parametrize("n_out", [4, 6])
@pytest.mark.parametrize("beams", ["equal", "unequal"])
@pytest.mark.parametrize("cfg", [dict(), dict(block=64, ept=3), dict(block=128, ept=1, lds=3)])
def test_friend_matches_oracle(n_out, beams, cfg):
    proc = W.process_friends(n_out)
    x, ch = _events(proc, 20011, 31 + n_out)
    E1 = E2 = None
    if beams == "unequal":
        E1, E2 = 0.6 * proc.sqrts, proc.sqrts ** 2 / (4 * 0.6 * proc.sqrts)
    print("friend parity", n_out, beams, cfg, flush=True)
    out, _, _ = run_gpu(proc, x, ch, E1=E1, E2=E2, **cfg)
    e = (None, None) if E1 is None else (np.full(len(ch), E1), np.full(len(ch), E2))
    assert_parity(forward_check(proc, x, ch, out, E1=e[0], E2=e[1]))


def test_root_friend_along_plus_z_equals_plain():
    proc = W.process_c1()
    x, ch = _events(proc, 5003, 2)
    plain, _, _ = run_gpu(proc, x, ch)
    q = proc.subset(range(proc.n_channels))
    q.friends = np.zeros_like(q.d1)
    q.friends[:, proc.n_prt_out - 1] = 1 << (proc.n_out + 1)
    fr, _, _ = run_gpu(q, x, ch)
    # every event is compared, including the non-ok ones whose momenta may be NaN
    for k in ("p", "factor", "volume", "ok"):
        assert np.array_equal(plain[k], fr[k], equal_nan=k != "ok"), k


def test_friend_trees_reproduce_whizard_momenta():
    """The 149 events of samples/bwbw.ref whose trees have friends, with the friend assignment
    that Whizard's own factors single out (test_whizard_friends.py), through the kernel."""
    import os
    from refio import read_ref
    from test_whizard_friends import FRIENDS
    proc, ev = read_ref(os.path.join(os.path.dirname(__file__), "golden", "bwbw_friends.ref"))
    proc.friends = np.zeros_like(proc.d1)
    for c, (b, which) in FRIENDS.items():
        proc.friends[c, b - 1] = 1 << (proc.n_out + which - 1)
    out, _, _ = run_gpu(proc, ev["x"], ev["channel"])
    assert out["ok"].all()
    dev = np.abs(out["p"] - ev["p"][:, proc.n_in:, :]).max() / proc.sqrts
    assert dev < 1e-6, dev
