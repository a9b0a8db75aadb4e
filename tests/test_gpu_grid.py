"""Adaptive grids on the GPU against the oracle's restatement: the randoms drawn from the grid
(Philox4x32-10 and the inverse distribution function) bit for bit, the weight update exactly,
and a short adaptation loop."""
import numpy as np
import pytest

import oracle
from gpu_phs_amd import workloads as W

pytestmark = pytest.mark.gpu


def _plan(proc, counts, **cfg):
    import torch
    import gpu_phs_amd as G
    ph = G.ProcessHandle(proc)
    segs, bufs = [], []
    for c, n in enumerate(counts):
        if n == 0:
            continue
        x = torch.zeros((n, proc.n_x), dtype=torch.float64, device="cuda")
        b = G.EventBuffers.allocate(x, proc.n_out)
        segs.append((ph, c, b))
        bufs.append(b)
    return ph, G.Plan(segs, **cfg), bufs, [c for c, n in enumerate(counts) if n]


@pytest.mark.parametrize("cfg", [dict(), dict(block=64, ept=3)])
def test_generate_matches_oracle(cfg):
    import torch
    import gpu_phs_amd as G
    proc = W.process_c1()
    ph, plan, bufs, chans = _plan(proc, [3001, 0, 777], **cfg)
    grid = G.Grid(ph)
    rng = np.random.default_rng(9)
    w = rng.random(grid.shape) + 0.01
    w[0, 2, 4] = 0.0
    grid.set_weights(w)
    seed = 0x1234_5678_9ABC
    grid.generate(plan, seed)
    torch.cuda.synchronize()
    F = oracle.grid_cdf(w)
    for s, (b, c) in enumerate(zip(bufs, chans)):
        got = b.x.cpu().numpy()
        want = oracle.grid_sample(F[c], oracle.grid_uniforms(seed, s, got.shape[0], proc.n_x))
        assert np.array_equal(got, want)


def test_update_matches_oracle_and_adapts():
    import torch
    import gpu_phs_amd as G
    proc = W.process_c1()
    counts = [20000, 15000, 5000]
    ph, plan, bufs, chans = _plan(proc, counts)
    grid = G.Grid(ph)
    w = grid.weights()
    assert np.allclose(w, 0.1)
    for it in range(3):
        grid.generate(plan, 1000 + it)
        plan.run()
        grid.update(plan)
        torch.cuda.synchronize()
        x = np.concatenate([b.x.cpu().numpy() for b in bufs])
        ok = np.concatenate([b.ok.cpu().numpy().astype(bool) for b in bufs])
        ch = np.repeat(chans, [b.x.shape[0] for b in bufs])
        want = oracle.grid_update(w, x, ok, ch)
        w = grid.weights()
        assert np.array_equal(w, want)
    # the weights now follow where accepted events lie: every dimension sums to one
    assert np.allclose(w.sum(axis=2), 1.0)
