"""GPU parity of the inverse mapping (all-channel factors and randoms) against the oracle, on
the momenta the GPU forward pass produced, and against Whizard's own all-channel factors."""
import os

import numpy as np
import pytest

import oracle
from gpu_phs_amd import workloads as W
from refio import read_ref
from test_gpu_forward import load_golden, run_gpu

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def inverse_stats(proc, f_gpu, x_gpu, f_ref, x_ref, own, use):
    """Relative factor and absolute random deviations over (event, channel) pairs other than
    the generating one, for the events in `use`; the phi randoms compare modulo 1."""
    n = f_gpu.shape[0]
    other = np.ones((n, proc.n_channels), bool)
    other[np.arange(n), own] = False
    other &= use[:, None]
    fr = f_ref[other]
    df = np.abs(f_gpu[other] - fr) / np.maximum(np.abs(fr), 1e-300)
    df[(fr == 0) & (f_gpu[other] == 0)] = 0
    xg = np.transpose(x_gpu, (1, 0, 2))[other]
    xr = np.transpose(x_ref, (1, 0, 2))[other]
    dx = np.abs(xg - xr)
    phi = [proc.n_out - 2 + 2 * k for k in range(proc.n_out - 1)]
    dx[:, phi] = np.minimum(dx[:, phi], 1 - dx[:, phi])
    return df, dx.max(axis=1)


def sensitivity(proc, p, f_ref, x_ref, own, use, E1=None, E2=None, samples=4, seed=0):
    """How far the oracle's own result moves when every momentum component is changed by one
    ulp (random signs, max over `samples` draws): a measured condition of each pair."""
    rng = np.random.default_rng(seed)
    sf = sx = None
    for _ in range(samples):
        pp = p * (1 + 2.0 ** -52 * rng.choice([-1.0, 1.0], size=p.shape))
        f2, x2 = oracle.inverse(proc, pp, E1=E1, E2=E2)
        a, b = inverse_stats(proc, f2, x2, f_ref, x_ref, own, use)
        sf = a if sf is None else np.maximum(sf, a)
        sx = b if sx is None else np.maximum(sx, b)
    return sf, sx


def assert_inverse_parity(df, dx, sf, sx):
    """To the last bits: the kernel adds final-state momenta daughter by daughter as the oracle
    does and evaluates the same formulas, so the factors agree to a few ulps and the randoms
    to about 1e-13 (its own atan2, log and sqrt against the host's).  A pair may exceed
    1e-13 / 1e-11 only where changing every momentum component by one ulp moves the oracle's
    own result by at least 1/64 of the deviation (nearly collinear massless pairs, peaked maps
    near cos theta = +-1)."""
    print(f"pairs {df.size}: factor rel max {df.max():.2e} q999 {np.quantile(df, 0.999):.2e}; "
          f"x max {dx.max():.2e} q999 {np.quantile(dx, 0.999):.2e}; "
          f"ill-conditioned (> 1e-13 / 1e-11) {(df > 1e-13).sum()} / {(dx > 1e-11).sum()}")
    assert np.all(np.isfinite(df)) and np.all(np.isfinite(dx))
    assert np.median(df) < 1e-15 and np.median(dx) < 1e-15
    assert np.quantile(df, 0.999) < 1e-14 and np.quantile(dx, 0.999) < 1e-13
    assert np.all((df <= 1e-13) | (df <= 64 * sf)) and np.all((dx <= 1e-11) | (dx <= 64 * sx))


PROCS = {"c1": W.process_c1, "c2": W.process_c2, "2to6": W.process_2to6,
         "friends4": lambda: W.process_friends(4), "friends6": lambda: W.process_friends(6),
         "edge": lambda: load_golden(os.path.join(GOLDEN, "golden_edge.npz"))[0]}


@pytest.mark.parametrize("name", sorted(PROCS))
@pytest.mark.parametrize("block", [0, 64, 128])
def test_inverse_matches_oracle(name, block):
    proc = PROCS[name]()
    rng = np.random.default_rng(17)
    n = 4001 if proc.n_channels < 50 else 1501
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    x = rng.random((n, proc.n_x))
    print("inverse parity", name, "block", block, flush=True)
    out, plan, bufs = run_gpu(proc, x, ch)
    f_gpu, x_gpu = plan.inverse(block=block)
    f_gpu, x_gpu = f_gpu.cpu().numpy(), x_gpu.cpu().numpy()
    # rows are the plan's segment order: events sorted by channel
    order = np.argsort(ch, kind="stable")
    p_sorted = out["p"][order]
    ok = out["ok"][order]
    own = ch[order]
    p_ok = np.where(ok[:, None, None], p_sorted, 0.0)
    f_ref, x_ref = oracle.inverse(proc, p_ok)
    # the generating channel's slot: the forward factor and the randoms themselves
    assert np.array_equal(f_gpu[np.arange(n), own], out["factor"][order], equal_nan=True)
    assert np.array_equal(x_gpu[own, np.arange(n)], x[order])
    df, dx = inverse_stats(proc, f_gpu, x_gpu, f_ref, x_ref, own, ok)
    assert_inverse_parity(df, dx, *sensitivity(proc, p_ok, f_ref, x_ref, own, ok))


@pytest.mark.parametrize("name", ["c1", "c2", "2to6", "large8", "edge", "beams_c2"])
def test_shared_roots_identical_to_per_channel(name):
    """The inverse that evaluates each root once per group of channels sharing it (and walks
    the trees of up to four final-state particles with kernels specialised per shape) against
    the kernel that inverts every channel on its own: the same operations on the same numbers,
    so the same bits."""
    from test_gpu_beams import _beams, run_beams
    proc = W.process_c2() if name == "beams_c2" else (W.process_large(8) if name == "large8" else PROCS[name]())
    rng = np.random.default_rng(29)
    n = 3001 if proc.n_channels < 50 else 1201
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    x = rng.random((n, proc.n_x))
    print("shared roots vs per channel", name, flush=True)
    if name.startswith("beams"):
        out, plan, _, _ = run_beams(proc, x, ch, *_beams(proc, n, rng))
    else:
        out, plan, _ = run_gpu(proc, x, ch)
    f1, x1 = (t.cpu().numpy() for t in plan.inverse())
    plan.set_specialize(False)
    f2, x2 = (t.cpu().numpy() for t in plan.inverse())
    assert np.array_equal(f1, f2, equal_nan=True)
    assert np.array_equal(x1, x2, equal_nan=True)


def test_inverse_unequal_beams():
    proc = W.process_friends(4)
    E1, E2 = 650.0, 1000.0 ** 2 / (4 * 650.0)
    rng = np.random.default_rng(3)
    n = 3001
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    x = rng.random((n, proc.n_x))
    out, plan, _ = run_gpu(proc, x, ch, E1=E1, E2=E2)
    f_gpu, x_gpu = (t.cpu().numpy() for t in plan.inverse())
    order = np.argsort(ch, kind="stable")
    ok = out["ok"][order]
    p_ok = np.where(ok[:, None, None], out["p"][order], 0.0)
    e1, e2 = np.full(n, E1), np.full(n, E2)
    f_ref, x_ref = oracle.inverse(proc, p_ok, E1=e1, E2=e2)
    df, dx = inverse_stats(proc, f_gpu, x_gpu, f_ref, x_ref, ch[order], ok)
    assert_inverse_parity(df, dx, *sensitivity(proc, p_ok, f_ref, x_ref, ch[order], ok, E1=e1, E2=e2))


@pytest.mark.parametrize("name,atol", [("ttH_head.ref", 5.01e-5), ("bwbw_head.ref", 5.01e-8),
                                       ("bwbw_friends.ref", 5.01e-8)])
def test_inverse_factors_match_whizard_on_gpu(name, atol):
    """Whizard's momenta in the segments' momentum buffers, every channel's factor against
    Whizard's all_factors (friend trees with the assignment of test_whizard_friends.py)."""
    import torch
    proc, ev = read_ref(os.path.join(GOLDEN, name))
    if name == "bwbw_friends.ref":
        from test_whizard_friends import FRIENDS
        proc.friends = np.zeros_like(proc.d1)
        for c, (b, which) in FRIENDS.items():
            proc.friends[c, b - 1] = 1 << (proc.n_out + which - 1)
        cols = np.ones(proc.n_channels, bool)
    else:
        cols = ev["hasfriends"] == 0
    ch = ev["channel"]
    out, plan, bufs = run_gpu(proc, ev["x"], ch)
    order = np.argsort(ch, kind="stable")
    pw = ev["p"][order][:, proc.n_in:, :]
    start = 0
    for b in bufs:
        m = b.p.shape[0]
        b.p.copy_(torch.from_numpy(np.ascontiguousarray(pw[start:start + m])))
        start += m
    f_gpu = plan.inverse()[0].cpu().numpy()
    fa = np.array(ev["factors"])[order]
    own = ch[order]
    other = np.ones_like(fa, bool)
    other[np.arange(len(own)), own] = False
    other &= cols[None, :]
    assert np.abs(f_gpu[other] - fa[other]).max() < atol
