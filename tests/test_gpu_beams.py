"""Per-event beam energies (phs.cu:1313-1317): every event its own sqrt(s) = 2 sqrt(E1 E2), lab
boost and sqrt(s)-dependent map constants, on the GPU, against the oracle, which re-derives the
constants for every event's sqrt(s) (phs_cpu.cpp:452-531)."""
import numpy as np
import pytest

import oracle
from gpu_phs_amd import workloads as W
from parity import assert_parity, forward_check
from test_gpu_forward import load_golden
from test_gpu_inverse import assert_inverse_parity, inverse_stats, sensitivity

pytestmark = pytest.mark.gpu


def _beams(proc, n, rng):
    """Beams around the process's sqrt(s): both energies vary, so do sqrt(s) and the lab boost."""
    h = proc.sqrts / 2
    return h * rng.uniform(0.8, 1.3, n), h * rng.uniform(0.8, 1.3, n)


def run_beams(proc, x, ch, E1, E2, specialize=True, **cfg):
    """specialize: the trees of up to four final-state particles on their shape's kernel (every
    shape its own launch, however few its events), else every tree on the generic kernel."""
    import torch
    import gpu_phs_amd as G
    ph = G.ProcessHandle(proc)
    order = np.argsort(ch, kind="stable")
    segs, bufs, beams = [], [], []
    start = 0
    xs = torch.from_numpy(np.ascontiguousarray(x[order])).cuda()
    e1s = torch.from_numpy(np.ascontiguousarray(E1[order])).cuda()
    e2s = torch.from_numpy(np.ascontiguousarray(E2[order])).cuda()
    chs = ch[order]
    for c in range(proc.n_channels):
        k = int((chs == c).sum())
        if k == 0:
            continue
        b = G.EventBuffers.allocate(xs[start:start + k], proc.n_out)
        segs.append((ph, c, b))
        bufs.append(b)
        beams.append((e1s[start:start + k], e2s[start:start + k]))
        start += k
    plan = G.Plan(segs, **cfg)
    plan.set_merge_below(0)
    if not specialize:
        plan.set_specialize(False)
    for i, (a, b) in enumerate(beams):
        plan.set_event_beams(i, a, b)
    plan.run()
    torch.cuda.synchronize()
    cat = lambda k: torch.cat([getattr(b, k) for b in bufs]).cpu().numpy()
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    out = dict(p=cat("p")[inv], factor=cat("factor")[inv], volume=cat("volume")[inv], ok=cat("ok")[inv].astype(bool))
    return out, plan, bufs, (e1s, e2s)


PROCS = {"c1": W.process_c1, "c2": W.process_c2, "2to6": W.process_2to6,
         "friends4": lambda: W.process_friends(4),
         "edge": lambda: load_golden(__import__("os").path.join(__import__("os").path.dirname(__file__), "golden",
                                                                "golden_edge.npz"))[0]}


@pytest.mark.parametrize("name", sorted(PROCS))
def test_per_event_beams_match_oracle(name):
    proc = PROCS[name]()
    rng = np.random.default_rng(41)
    n = 6001
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    x = rng.random((n, proc.n_x))
    E1, E2 = _beams(proc, n, rng)
    print("per-event beams", name, flush=True)
    out, plan, _, _ = run_beams(proc, x, ch, E1, E2)
    assert_parity(forward_check(proc, x, ch, out, E1=E1, E2=E2))


@pytest.mark.parametrize("name", ["c1", "c2", "2to6", "friends4"])
def test_per_event_beams_inverse(name):
    proc = PROCS[name]()
    rng = np.random.default_rng(43)
    n = 3001 if proc.n_channels < 50 else 1201
    print("per-event beams inverse", name, flush=True)
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    x = rng.random((n, proc.n_x))
    E1, E2 = _beams(proc, n, rng)
    out, plan, _, _ = run_beams(proc, x, ch, E1, E2)
    f_gpu, x_gpu = (t.cpu().numpy() for t in plan.inverse())
    order = np.argsort(ch, kind="stable")
    ok = out["ok"][order]
    p_ok = np.where(ok[:, None, None], out["p"][order], 0.0)
    e1, e2 = E1[order], E2[order]
    f_ref, x_ref = oracle.inverse(proc, p_ok, E1=e1, E2=e2)
    df, dx = inverse_stats(proc, f_gpu, x_gpu, f_ref, x_ref, ch[order], ok)
    assert_inverse_parity(df, dx, *sensitivity(proc, p_ok, f_ref, x_ref, ch[order], ok, E1=e1, E2=e2))


def test_per_event_beams_built_in_loss_is_rejected():
    """The built-in loss of the adjoint (loss_mode 1) assumes equal beams; with per-event beam
    energies the adjoint takes the caller's cotangents (loss_mode 0, test_gpu_adjoint.py)."""
    import torch
    import gpu_phs_amd as G
    proc = W.process_c1()
    rng = np.random.default_rng(5)
    n = 500
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    E1, E2 = _beams(proc, n, rng)
    _, plan, bufs, _ = run_beams(proc, rng.random((n, proc.n_x)), ch, E1, E2)
    with pytest.raises(G.GpuPhsError, match="per-event beam"):
        plan.adjoint([torch.zeros_like(b.x) for b in bufs], loss_mode=1)


@pytest.mark.parametrize("n_out", [2, 3, 4])
@pytest.mark.parametrize("cfg", [dict(), dict(block=64, ept=3, waves_per_eu=3), dict(block=128, ept=2, waves_per_eu=5),
                                 dict(block=128, ept=1)])
def test_per_event_beams_shape_kernels_identical_to_generic(n_out, cfg):
    """With per-event beams, the kernels specialised for the trees of 2-4 final-state particles
    give the generic kernel's bits: momenta, factors, volumes and ok flags of every event."""
    proc = W.process_shapes(n_out)
    rng = np.random.default_rng(40 + n_out)
    n = 20011
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    x = rng.random((n, proc.n_x))
    E1, E2 = _beams(proc, n, rng)
    print("per-event beams shapes", n_out, cfg, flush=True)
    spec, _, _, _ = run_beams(proc, x, ch, E1, E2, **cfg)
    gen, _, _, _ = run_beams(proc, x, ch, E1, E2, specialize=False, **cfg)
    for k in ("p", "factor", "volume", "ok"):
        assert np.array_equal(spec[k], gen[k], equal_nan=k != "ok"), (k, cfg)
    assert spec["ok"].sum() > 1000
    assert_parity(forward_check(proc, x, ch, spec, E1=E1, E2=E2))
