"""The friend axis in the oracle (CPU): branches whose polar axis is an incoming beam.

The reference has this only on the GPU (phs.cu:853-910), with a rotation matrix that is not
orthogonal (L_rot_to_2nd, phs.cu:753-781), and no CPU counterpart or fixture, so its parity is
unpinned.  What fixes the construction instead:

* the momenta stay on shell and conserve the total momentum (the rotation is a rotation);
* the polar angle of a friend branch's first daughter, measured from the friend beam in the
  branch's rest frame, is the angle its cos-theta map gave;
* a friend on the root with f = +1 (beam along +z) is the plain composition, bit for bit;
* a friend on the root with f = -1 is the plain composition with theta -> pi - theta,
  phi -> -phi, and the root's internal daughters turned by a half turn about their axis.
"""
import numpy as np
import pytest

import oracle
from gpu_phs_amd import process as P
from gpu_phs_amd import workloads as W


def _events(proc, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, proc.n_x))
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    return x, ch


def _leaves(node):
    if isinstance(node, int):
        return [node]
    return _leaves(node[0]) + _leaves(node[1])


def _tree_of(proc, c):
    """Nested-tuple tree of channel c from the bincode arrays."""
    def node(b):
        if not proc.has_children[c, b - 1]:
            return int(np.log2(b))
        return (node(int(proc.d1[c, b - 1])), node(int(proc.d2[c, b - 1])))
    return node(proc.n_prt_out)


def _rest_frame(P4, q):
    """q in the rest frame of P4 (rows of four-vectors)."""
    m = np.sqrt(P4[:, 0] ** 2 - np.sum(P4[:, 1:] ** 2, axis=1))
    beta = P4[:, 1:] / P4[:, :1]
    g = P4[:, 0] / m
    b2 = np.sum(beta ** 2, axis=1)
    bq = np.sum(beta * q[:, 1:], axis=1)
    coef = np.where(b2 > 0, (g - 1) * bq / np.where(b2 > 0, b2, 1), 0.0) - g * q[:, 0]
    return q[:, 1:] + coef[:, None] * beta


@pytest.mark.parametrize("n_out", [4, 6])
@pytest.mark.parametrize("beams", ["equal", "unequal"])
def test_friend_conservation_and_mass_shell(n_out, beams):
    proc = W.process_friends(n_out)
    x, ch = _events(proc, 2000, 11 + n_out)
    E1 = E2 = None
    e1, e2 = proc.sqrts / 2, proc.sqrts / 2
    if beams == "unequal":
        e1, e2 = 0.7 * proc.sqrts, proc.sqrts ** 2 / (4 * 0.7 * proc.sqrts)
        E1, E2 = np.full(len(ch), e1), np.full(len(ch), e2)
    o = oracle.forward(proc, x, ch, E1=E1, E2=E2)
    ok = o["ok"]
    assert ok.sum() > 200
    p = o["p"][ok]
    tot = p.sum(axis=1)
    want = np.array([e1 + e2, 0.0, 0.0, e1 - e2])
    assert np.allclose(tot, want, rtol=0, atol=1e-9 * proc.sqrts)
    m2 = p[..., 0] ** 2 - np.sum(p[..., 1:] ** 2, axis=-1)
    flv = proc.flv_masses[2:] ** 2
    assert np.allclose(m2, np.broadcast_to(flv, m2.shape), rtol=0, atol=1e-8 * proc.sqrts ** 2)


@pytest.mark.parametrize("n_out", [4, 6])
@pytest.mark.parametrize("beams", ["equal", "unequal"])
def test_friend_polar_angle_is_measured_from_the_beam(n_out, beams):
    proc = W.process_friends(n_out)
    x, ch = _events(proc, 3000, 23 + n_out)
    e1 = e2 = proc.sqrts / 2
    E1 = E2 = None
    if beams == "unequal":
        e1, e2 = 0.55 * proc.sqrts, proc.sqrts ** 2 / (4 * 0.55 * proc.sqrts)
        E1, E2 = np.full(len(ch), e1), np.full(len(ch), e2)
    o = oracle.forward(proc, x, ch, E1=E1, E2=E2)
    s_f = 4 * e1 * e2
    checked = 0
    for c in range(proc.n_channels):
        tree = _tree_of(proc, c)
        for idx, node in enumerate(P.internal_nodes(tree)):
            b = P._bincode(node)
            fr = int(proc.friends[c, b - 1])
            if fr == 0:
                continue
            f = -1 if fr == 1 << proc.n_out else 1
            sel = (ch == c) & o["ok"]
            p = o["p"][sel]
            k1 = p[:, _leaves(node[0])].sum(axis=1)
            k2 = p[:, _leaves(node[1])].sum(axis=1)
            Ef = e1 if f == 1 else e2
            beam = np.tile([Ef, 0.0, 0.0, f * Ef], (k1.shape[0], 1))
            q = _rest_frame(k1 + k2, k1)
            kb = _rest_frame(k1 + k2, beam)
            cos = np.sum(q * kb, axis=1) / np.linalg.norm(q, axis=1) / np.linalg.norm(kb, axis=1)
            xct = x[sel, proc.n_out - 2 + 2 * idx + 1]
            mid = int(proc.map_ids[c, b - 1])
            m = float(proc.map_masses[c, b - 1])
            b0 = m * m / s_f
            bb = np.array([b0, np.log((b0 + 1) / b0) if b0 > 0 else 0.0, 0.0])
            want = np.array([oracle.map_ct(mid, float(v), bb)[0] for v in xct])
            assert np.allclose(cos, want, rtol=0, atol=1e-8), (c, node)
            checked += 1
    assert checked == int(np.count_nonzero(proc.friends))


def _root_friend(proc, which):
    q = proc.subset(range(proc.n_channels))
    q.friends = np.zeros_like(q.d1)
    q.friends[:, proc.n_prt_out - 1] = 1 << (proc.n_out + which - 1)
    return q


@pytest.mark.parametrize("make", [W.process_c1, W.process_c2, W.process_2to6])
def test_root_friend_along_plus_z_is_the_plain_composition(make):
    proc = make()
    x, ch = _events(proc, 1000, 5)
    plain = oracle.forward(proc, x, ch)
    fr = oracle.forward(_root_friend(proc, 2), x, ch)
    for k in ("p", "factor", "volume", "ok"):
        assert np.array_equal(plain[k], fr[k]), k


def test_root_friend_along_minus_z_turns_the_frame():
    # flat cos-theta maps at the root: theta -> pi - theta is x_ct -> 1 - x_ct exactly
    proc = W.process_c1()
    proc.map_ids[:, proc.n_prt_out - 1] = P.MAP_NO
    x, ch = _events(proc, 1000, 9)
    fr = oracle.forward(_root_friend(proc, 1), x, ch)
    x2 = x.copy()
    n0 = proc.n_out - 2            # the root's phi random; cos theta follows it
    x2[:, n0] = 1 - x[:, n0]
    x2[:, n0 + 1] = 1 - x[:, n0 + 1]
    # the root's internal daughters: phi + pi
    for c in range(proc.n_channels):
        tree = _tree_of(proc, c)
        inner = P.internal_nodes(tree)
        for i, node in enumerate(inner[1:], start=1):
            if node in (tree[0], tree[1]):
                j = n0 + 2 * i
                sel = ch == c
                x2[sel, j] = np.mod(x[sel, j] + 0.5, 1.0)
    plain = oracle.forward(proc, x2, ch)
    assert np.array_equal(plain["ok"], fr["ok"])
    ok = fr["ok"]
    assert np.allclose(fr["p"][ok], plain["p"][ok], rtol=0, atol=1e-9 * proc.sqrts)
    assert np.allclose(fr["factor"][ok], plain["factor"][ok], rtol=1e-12, atol=0)


def test_friend_sign_from_the_friends_array():
    # 2^n_out (incoming particle 1) gives -1, 2^(n_out + 1) gives +1 (phs.cu:276-284); any
    # other entry is no friend
    proc = W.process_c1()
    x, ch = _events(proc, 400, 3)
    plain = oracle.forward(proc, x, ch)
    q = proc.subset(range(proc.n_channels))
    q.friends = np.zeros_like(q.d1)
    q.friends[:, proc.n_prt_out - 1] = 3    # not an incoming bincode
    assert np.array_equal(oracle.forward(q, x, ch)["p"], plain["p"])
    minus = oracle.forward(_root_friend(proc, 1), x, ch)
    ok = minus["ok"]
    assert not np.allclose(minus["p"][ok], plain["p"][ok])


def test_forward_dual_follows_the_friend_axis():
    """The dual-number forward takes the friend axis: its primal values are the friend
    forward's, bit for bit, and differ from those of the same trees without friends."""
    proc = W.process_friends(4)
    x, ch = _events(proc, 400, 1)
    d = oracle.forward_dual(proc, x, ch)
    f = oracle.forward(proc, x, ch)
    ok = f["ok"]
    assert np.array_equal(d["ok"], ok)
    assert np.array_equal(d["p"][ok], f["p"][ok]) and np.array_equal(d["factor"][ok], f["factor"][ok])
    q = proc.subset(range(proc.n_channels))
    q.friends = None
    plain = oracle.forward_dual(q, x, ch)
    both = ok & plain["ok"]
    assert not np.allclose(plain["p"][both], d["p"][both])
