"""The forward parity criterion (tests/parity.py) on the CPU: another correct FP64 evaluation of
the reference algorithm passes it, and planted mistakes -- including ones confined to the
ill-conditioned events that the criterion treats specially -- fail it.  The stand-in for the
HIP path is the oracle itself, run with its elementary functions moved by one ulp along a seed
that the sensitivity measurement does not use."""
import numpy as np
import pytest

import oracle
from gpu_phs_amd import workloads as W
from parity import assert_parity, forward_check

N = 6000


def _case(name):
    proc = {"c1": W.process_c1, "c2": W.process_c2, "2to6": W.process_2to6}[name]()
    rng = np.random.default_rng(101)
    ch = rng.integers(0, proc.n_channels, N).astype(np.int32)
    x = rng.random((N, proc.n_x))
    ref = oracle.forward(proc, x, ch, want_msq=True)
    other = oracle.forward(proc, x, ch, want_msq=True, perturb=0xC0FFEE)
    return proc, x, ch, ref, other


def _fails(proc, x, ch, out, ref):
    with pytest.raises(AssertionError):
        assert_parity(forward_check(proc, x, ch, out, ref=ref))


@pytest.mark.parametrize("name", ["c1", "c2", "2to6"])
def test_other_correct_evaluation_passes(name):
    proc, x, ch, ref, other = _case(name)
    out = {k: other[k] for k in ("p", "factor", "volume", "ok", "msq")}
    assert_parity(forward_check(proc, x, ch, out, ref=ref))


def test_error_on_ill_conditioned_events_fails():
    """1e-10 relative on the momenta of the 20 most sensitive ok events only."""
    proc, x, ch, ref, other = _case("c2")
    sens = oracle.sensitivity(proc, x, ch, base=ref)
    s = (sens["p"] / np.maximum(np.abs(ref["p"]), 1e-300)).reshape(N, -1).max(axis=1)
    s[~ref["ok"]] = -1
    idx = np.argsort(-s)[:20]
    out = dict(p=ref["p"].copy(), factor=ref["factor"], volume=ref["volume"], ok=ref["ok"])
    out["p"][idx] *= 1 + 1e-10
    _fails(proc, x, ch, out, ref)


def test_scattered_small_errors_fail():
    """3e-12 relative on 1% of the events: each below the cap, too many to be explained."""
    proc, x, ch, ref, _ = _case("c1")
    out = dict(p=ref["p"].copy(), factor=ref["factor"], volume=ref["volume"], ok=ref["ok"])
    idx = np.random.default_rng(3).choice(np.nonzero(ref["ok"])[0], N // 100, replace=False)
    out["p"][idx, :, 1] *= 1 + 3e-12
    _fails(proc, x, ch, out, ref)


def test_sign_error_on_one_event_fails():
    proc, x, ch, ref, _ = _case("2to6")
    out = dict(p=ref["p"].copy(), factor=ref["factor"], volume=ref["volume"], ok=ref["ok"])
    i = np.nonzero(ref["ok"])[0][5]
    out["p"][i, 2, 2] = -out["p"][i, 2, 2]
    _fails(proc, x, ch, out, ref)


def test_ok_flip_away_from_threshold_fails():
    proc, x, ch, ref, _ = _case("c1")
    out = dict(p=ref["p"], factor=ref["factor"], volume=ref["volume"], ok=ref["ok"].copy())
    i = np.nonzero(ref["ok"])[0][0]
    out["ok"][i] = False
    _fails(proc, x, ch, out, ref)


def test_factor_error_fails():
    proc, x, ch, ref, _ = _case("c2")
    out = dict(p=ref["p"], factor=ref["factor"] * (1 + 5e-12), volume=ref["volume"], ok=ref["ok"])
    _fails(proc, x, ch, out, ref)


def test_msq_error_fails():
    proc, x, ch, ref, other = _case("c1")
    out = {k: other[k] for k in ("p", "factor", "volume", "ok")}
    out["msq"] = other["msq"].copy()
    out["msq"][:, 2] *= 1 + 1e-11     # bincode 3: the (12) resonance of the first channel
    _fails(proc, x, ch, out, ref)


def test_sensitivity_is_zero_without_elementary_functions():
    """The 2->2 process with a flat polar map and phi = 0.25, 0.5: only exact or correctly
    rounded operations are left, and the perturbation must not move them beyond an ulp."""
    proc = W.process_c0()
    x = np.array([[0.25, 0.5], [0.5, 0.75]])
    ch = np.zeros(2, np.int32)
    base = oracle.forward(proc, x, ch)
    sens = oracle.sensitivity(proc, x, ch, base=base)
    assert (sens["p"] <= 4 * np.spacing(np.abs(base["p"]).max())).all()
    assert not sens["ok_flip"].any()
