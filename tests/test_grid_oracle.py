"""The grid restatement (CPU): Philox4x32-10 against its published known-answer vectors, the
inverse distribution function of the piecewise-constant density, and the reference's weight
update (rng.cpp:33-67) on hand-built events."""
import numpy as np

import oracle


def test_philox4x32_10_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    cases = [((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
             ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
             ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
              (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in cases:
        got = oracle.philox4x32_10(np.array([ctr], np.uint32), key)[0]
        assert tuple(int(v) for v in got) == want


def test_uniform_grid_is_the_identity():
    F = oracle.grid_cdf(np.full((3, 10), 0.1))
    u = oracle.grid_uniforms(12345, 0, 20000, 3)
    x = oracle.grid_sample(F, u)
    assert np.abs(x - u).max() < 4e-16


def test_grid_sample_follows_the_bin_weights():
    rng = np.random.default_rng(3)
    w = rng.random((2, 10)) + 0.05
    w[1, 3] = 0.0                      # an empty bin is never drawn
    F = oracle.grid_cdf(w)
    n = 400000
    x = oracle.grid_sample(F, oracle.grid_uniforms(7, 2, n, 2))
    assert (x >= 0).all() and (x < 1).all()
    for k in range(2):
        counts = np.bincount((x[:, k] * 10).astype(int), minlength=10)
        p = w[k] / w[k].sum()
        assert counts[p == 0].sum() == 0
        sel = p > 0
        chi2 = (((counts - n * p) ** 2)[sel] / (n * p[sel])).sum()
        assert chi2 < 40, chi2          # 9 degrees of freedom: P(chi2 > 40) ~ 1e-5
        # within a bin the density is flat
        b = int(np.argmax(p))
        inside = x[(x[:, k] * 10).astype(int) == b, k] * 10 - b
        assert abs(inside.mean() - 0.5) < 0.01


def test_grid_sample_is_monotone_in_u():
    w = np.array([[0.3, 0.0, 0.05, 0.1, 0.2, 0.05, 0.1, 0.1, 0.05, 0.05]])
    u = np.sort(oracle.grid_uniforms(1, 0, 5000, 1), axis=0)
    x = oracle.grid_sample(oracle.grid_cdf(w), u)
    assert (np.diff(x[:, 0]) >= 0).all()


def test_grid_update_known_answer():
    w = np.full((2, 2, 10), 0.1)
    x = np.array([[0.05, 0.95], [0.15, 0.95], [0.999, 0.0], [0.5, 0.5]])
    ok = np.array([True, True, True, False])
    ch = np.array([0, 0, 0, 1])
    out = oracle.grid_update(w, x, ok, ch)
    want0 = np.zeros(10); want0[[0, 1, 9]] = 1 / 3
    want1 = np.zeros(10); want1[9] = 2 / 3; want1[0] = 1 / 3
    assert np.array_equal(out[0, 0], want0) and np.array_equal(out[0, 1], want1)
    assert np.array_equal(out[1], w[1])   # no accepted event: unchanged
