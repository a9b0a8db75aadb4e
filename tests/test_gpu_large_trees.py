"""Trees with 8 and 10 final-state particles on the GPU against the oracle: save banks up to
three levels deep (in LDS), randoms in LDS (n_x = 20, 26), and in the inverse the daughter sum
programs of subtrees of up to seven particles.  Deep random trees are kinematically allowed
only for a small share of the randoms (every branch must be heavier than its daughters), so
the samples are large and the checks count the allowed events."""
import numpy as np
import pytest

import oracle
from gpu_phs_amd import workloads as W
from parity import assert_parity, forward_check
from test_gpu_forward import run_gpu
from test_gpu_inverse import assert_inverse_parity, inverse_stats, sensitivity

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_out", [8, 10])
@pytest.mark.parametrize("cfg", [{}, dict(block=128, ept=2, lds=1)], ids=["default", "b128"])
def test_large_trees_forward(n_out, cfg):
    proc = W.process_large(n_out)
    rng = np.random.default_rng(23 + n_out)
    n = 200_003
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    x = rng.random((n, proc.n_x))
    print("large trees", n_out, cfg, flush=True)
    out, _, _ = run_gpu(proc, x, ch, **cfg)
    r = forward_check(proc, x, ch, out)
    assert r["n_ok"] > 100
    assert_parity(r)


def test_large_trees_inverse():
    proc = W.process_large(8)
    rng = np.random.default_rng(29)
    n = 60_001
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    x = rng.random((n, proc.n_x))
    print("large trees inverse", flush=True)
    out, plan, _ = run_gpu(proc, x, ch)
    f_gpu, x_gpu = (t.cpu().numpy() for t in plan.inverse())
    order = np.argsort(ch, kind="stable")
    ok = out["ok"][order]
    assert ok.sum() > 300
    own = ch[order]
    p_ok = np.where(ok[:, None, None], out["p"][order], 0.0)
    assert np.array_equal(f_gpu[np.arange(n), own], out["factor"][order], equal_nan=True)
    f_ref, x_ref = oracle.inverse(proc, p_ok)
    df, dx = inverse_stats(proc, f_gpu, x_gpu, f_ref, x_ref, own, ok)
    assert_inverse_parity(df, dx, *sensitivity(proc, p_ok, f_ref, x_ref, own, ok))
