"""Pinning the oracle (CPU restatement of the reference's phs_cpu.cpp).

Each check is chosen so that a plausible restatement mistake fails it: the golden vectors
and the live reference build pin every expression bitwise; Whizard's own momenta in the
reference's sample files pin the conventions invariants cannot see (random order, daughter
signs, rotation order) -- the planted-mistake tests show that they do; closed forms,
conservation laws and finite differences pin the maps and the decay kinematics.
"""
import glob
import os

import numpy as np
import pytest

import oracle
from gpu_phs_amd import process as P
from gpu_phs_amd import workloads as W
from refio import read_ref

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
PROCS = {"c0": W.process_c0, "c1": W.process_c1, "c2": W.process_c2, "2to6": W.process_2to6,
         "2to8": lambda: W.process_large(8)}


def load_golden(path):
    z = np.load(path)
    proc = P.Process(int(z["n_in"]), int(z["n_out"]), z["d1"], z["d2"], z["has_children"], z["map_ids"],
                     z["map_masses"], z["map_widths"], z["flv_masses"], z["flv_widths"], float(z["sqrts"]))
    return proc, z


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "golden_*.npz"))),
                         ids=lambda p: os.path.basename(p)[7:-4])
def test_oracle_bitwise_equals_golden(path):
    """Golden vectors written by the compiled reference phs_cpu.cpp (tests/golden/make_golden.py)."""
    proc, z = load_golden(path)
    r = oracle.forward(proc, z["x"], z["channels"], want_msq=True)
    assert np.array_equal(r["ok"], z["ok"])
    ok = z["ok"]
    assert np.array_equal(r["p"][ok], z["p"][ok])
    assert np.array_equal(r["factor"][ok], z["factor"][ok])
    assert np.array_equal(r["volume"][ok], z["volume"][ok])
    assert np.array_equal(r["msq"], z["msq"], equal_nan=True)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference CPU build (oracle/_ref) not present")
@pytest.mark.parametrize("name", list(PROCS))
def test_oracle_bitwise_equals_reference_build(name):
    proc = PROCS[name]()
    rng = np.random.default_rng(2024)
    n = 5000
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    x = rng.random((n, proc.n_x))
    a = oracle.forward(proc, x, ch, want_msq=True)
    b = oracle.ref_forward(proc, x, ch, want_msq=True)
    assert np.array_equal(a["ok"], b["ok"])
    ok = a["ok"]
    for k in ("p", "factor", "volume"):
        assert np.array_equal(a[k][ok], b[k][ok]), k
    assert np.array_equal(a["msq"], b["msq"], equal_nan=True)


@pytest.mark.skipif(not oracle.ref_available(), reason="reference CPU build (oracle/_ref) not present")
def test_reference_timing_threads_cover_every_event():
    """The all-cores CPU baseline splits the events over threads; its ok count is the
    single-thread loop's and the reference build's forward pass's."""
    proc = PROCS["c1"]()
    rng = np.random.default_rng(5)
    n = 20011
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    x = rng.random((n, proc.n_x))
    n_ok = int(oracle.ref_forward(proc, x, ch)["ok"].sum())
    assert oracle.ref_time(proc, x, ch)[1] == n_ok
    for threads in (1, 3, 8):
        assert oracle.ref_time_threads(proc, x, ch, threads)[1] == n_ok
    with pytest.raises(RuntimeError):
        oracle.ref_time_threads(proc, x, ch, 0)


def _whizard_check(path, perm=None, only_nofriends=True):
    proc, ev = read_ref(path)
    sel = np.ones(len(ev["channel"]), bool)
    if only_nofriends:
        sel &= ev["hasfriends"][ev["channel"]] == 0
    x = ev["x"][sel]
    if perm is not None:
        x = x[:, perm]
    r = oracle.forward(proc, x, ev["channel"][sel])
    p_w = ev["p"][sel][:, proc.n_in:, :]
    scale = proc.sqrts
    dev = np.abs(r["p"] - p_w).max(axis=(1, 2)) / scale
    return r, ev, sel, dev


def test_whizard_momenta_tth():
    """samples/ttH.ref (2->3, three channels): Whizard's momenta from its 9-decimal randoms."""
    r, ev, sel, dev = _whizard_check(os.path.join(GOLDEN, "ttH_head.ref"))
    assert r["ok"].all() and ev["ok"][sel].all()
    assert dev.max() < 1e-6, dev.max()
    assert np.allclose(r["volume"], ev["volume"][sel], rtol=1e-12)


def test_whizard_momenta_bwbw():
    """samples/bwbw.ref (2->4, 61 channels incl. BW and t-channel maps); channels whose tree has
    friends are skipped here (the reference's CPU algorithm has no friend axis and the file no
    per-branch friend data; test_whizard_friends.py covers them)."""
    r, ev, sel, dev = _whizard_check(os.path.join(GOLDEN, "bwbw_head.ref"))
    assert sel.sum() > 50
    assert r["ok"].all()
    assert dev.max() < 1e-6, dev.max()
    assert np.allclose(r["volume"], ev["volume"][sel], rtol=1e-12)


@pytest.mark.parametrize("perm", [[1, 0, 2, 3, 4], [0, 1, 3, 2, 4], [2, 3, 4, 0, 1]],
                         ids=["msq-order", "phi-costheta-swapped", "angles-first"])
def test_whizard_catches_random_order_mistakes(perm):
    """Planted mistake (in the inputs of the oracle, on the CPU): a wrong random-number order
    must fail the Whizard comparison."""
    _, _, _, dev = _whizard_check(os.path.join(GOLDEN, "ttH_head.ref"), perm=perm)
    assert dev.max() > 1e-3


def test_closed_form_2to2():
    """e+e- -> t tbar at sqrt(s) = 500: |p| = sqrt(lambda(s, m^2, m^2)) / (2 sqrt s),
    factor = sqrt(lambda)/s, volume = 1 / (4 (2 pi)^5), back-to-back momenta."""
    proc = W.process_c0()
    rng = np.random.default_rng(5)
    x = rng.random((1000, 2))
    r = oracle.forward(proc, x, np.zeros(1000, np.int32))
    s, m = 500.0 ** 2, 173.0
    lam = (s - 2 * m * m) ** 2 - 4 * m ** 4
    pabs = np.linalg.norm(r["p"][:, 0, 1:], axis=1)
    assert np.allclose(pabs, np.sqrt(lam) / (2 * 500.0), rtol=1e-13)
    assert np.allclose(pabs, 180.4743749123404, rtol=1e-13)
    assert np.allclose(r["factor"], np.sqrt(lam) / s, rtol=1e-13)
    assert np.allclose(r["factor"], 0.7218974996493616, rtol=1e-13)
    assert np.allclose(r["volume"], 1 / (4 * (2 * np.pi) ** 5), rtol=1e-13)
    assert np.allclose(r["p"][:, 0, 1:], -r["p"][:, 1, 1:], rtol=0, atol=1e-10)
    # phi = 2 pi x0, cos theta = 2 x1 - 1 for MAP_NO (x0 first, then x1)
    phi = np.arctan2(r["p"][:, 0, 2], r["p"][:, 0, 1]) % (2 * np.pi)
    assert np.allclose(phi, 2 * np.pi * x[:, 0], atol=1e-9)
    assert np.allclose(r["p"][:, 0, 3] / pabs, 2 * x[:, 1] - 1, atol=1e-12)


@pytest.mark.parametrize("name", list(PROCS))
def test_conservation_on_shell_volume(name):
    proc = PROCS[name]()
    rng = np.random.default_rng(9)
    n = 3000
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    r = oracle.forward(proc, rng.random((n, proc.n_x)), ch)
    ok = r["ok"]
    p = r["p"][ok]
    tot = p.sum(axis=1)
    assert np.allclose(tot[:, 0], proc.sqrts, rtol=1e-11)
    assert np.allclose(tot[:, 1:], 0, atol=1e-9 * proc.sqrts)
    m2 = p[..., 0] ** 2 - (p[..., 1:] ** 2).sum(-1)
    mass = proc.flv_masses[proc.n_in:]
    assert np.allclose(m2, mass ** 2, atol=1e-8 * proc.sqrts ** 2)
    s, k = proc.sqrts ** 2, proc.n_out
    vol = s ** (k - 2) / (4 ** (k - 1) * (2 * np.pi) ** (2 * k + 1))
    assert np.allclose(r["volume"][ok], vol, rtol=1e-13)


def _map_probe(map_id, m=91.2, w=2.5, sqrts=800.0):
    proc = P.make_process(2, 3, [(0, (1, 2))], [{(1, 2): (map_id, m, w)}], [0.000511, 0.000511, 4.8, 0.0, 1.0],
                          sqrts)
    a, b = oracle.constants(proc)
    ms = oracle.mass_sums(proc)
    bidx = 6 - 1  # bincode (1,2) -> 2|4 = 6
    m_min = ms[0, bidx]
    m_max = sqrts - ms[0, 6] + m_min
    return a[0, bidx], b[0, bidx], m_min ** 2, m_max ** 2, sqrts ** 2


@pytest.mark.parametrize("map_id", [P.MAP_NO, P.MAP_SCHANNEL, P.MAP_TCHANNEL, P.MAP_UCHANNEL, P.MAP_RADIATION,
                                    P.MAP_COLLINEAR, P.MAP_INFRARED, P.MAP_STEP_E, P.MAP_STEP_H])
def test_msq_map_endpoints_and_jacobian(map_id):
    """msq(0) = msq_min, msq(1) = msq_max and f_map = (d msq / dx) / s (central differences).
    STEP_E reaches msq_max through log(z1 - a1) with z1 - a1 = exp(-1/a2), a cancellation the
    reference's map has; a wide resonance (a2 = 0.044) keeps that end point representable."""
    m, w = (350.0, 40.0) if map_id == P.MAP_STEP_E else (91.2, 2.5)
    a, _, lo, hi, s = _map_probe(map_id, m, w)
    f = lambda x: oracle.map_msq(map_id, x, s, m, w, lo, hi, a)
    assert np.isclose(f(0.0)[0], lo, rtol=1e-9, atol=1e-9 * hi)
    assert np.isclose(f(1.0)[0], hi, rtol=1e-9 if map_id != P.MAP_STEP_E else 1e-5)
    h = 1e-6
    for x in (0.1, 0.3, 0.45, 0.55, 0.7, 0.9):
        d = (f(x + h)[0] - f(x - h)[0]) / (2 * h) / s
        assert np.isclose(f(x)[1], d, rtol=1e-6), (map_id, x)


@pytest.mark.parametrize("map_id", [P.MAP_NO, P.MAP_TCHANNEL, P.MAP_COLLINEAR])
def test_ct_map_endpoints_and_jacobian(map_id):
    """cos theta(0) = -1, cos theta(1) = 1, f_ct = (d ct / dx) / 2, st = sqrt(1 - ct^2)."""
    _, b, _, _, _ = _map_probe(map_id, m=30.0)
    f = lambda x: oracle.map_ct(map_id, x, b)
    assert np.isclose(f(0.0)[0], -1, atol=1e-12) and np.isclose(f(1.0)[0], 1, atol=1e-12)
    h = 1e-6
    for x in (0.05, 0.3, 0.45, 0.55, 0.8, 0.95):
        ct, st, fc = f(x)
        assert np.isclose(st, np.sqrt(1 - ct * ct), rtol=1e-12)
        d = (f(x + h)[0] - f(x - h)[0]) / (2 * h) / 2
        assert np.isclose(fc, d, rtol=1e-6), (map_id, x)


def _dual_against_differences(proc, n, seed, beams=False):
    """The oracle's dual numbers against central differences of the oracle itself: d p / d x,
    d factor / d x and d (p, factor) / d (M, Gamma) of every mapped branch (Breit-Wigner masses
    and widths, and the masses of the t-channel maps, which enter through the cos-theta
    constants b0 = M^2 / s), with fixed or per-event beam energies."""
    rng = np.random.default_rng(seed)
    ch = rng.integers(0, proc.n_channels, n).astype(np.int32)
    x = 0.05 + 0.9 * rng.random((n, proc.n_x))
    E1 = E2 = None
    if beams:
        h2 = proc.sqrts / 2
        E1, E2 = h2 * rng.uniform(0.8, 1.3, n), h2 * rng.uniform(0.8, 1.3, n)
    K = proc.n_x + 2
    xt = np.zeros((n, proc.n_x, K))
    for k in range(proc.n_x):
        xt[:, k, k] = 1
    mt = np.zeros((proc.n_channels, proc.n_prt_out, K))
    wt = np.zeros_like(mt)
    mapped = (proc.map_ids == P.MAP_SCHANNEL) | (proc.map_ids == P.MAP_TCHANNEL)
    mt[mapped, proc.n_x] = 1
    wt[proc.map_ids == P.MAP_SCHANNEL, proc.n_x + 1] = 1
    d = oracle.forward_dual(proc, x, ch, xt, mt, wt, K, E1=E1, E2=E2)
    ok = d["ok"]
    assert ok.sum() > n // 4
    fwd = lambda q, xx: oracle.forward(q, xx, ch, E1=E1, E2=E2)
    base = fwd(proc, x)
    assert np.array_equal(base["ok"], ok)
    assert np.array_equal(base["p"][ok], d["p"][ok]) and np.array_equal(base["factor"][ok], d["factor"][ok])
    h = 1e-6
    for k in range(proc.n_x):
        xp, xm = x.copy(), x.copy()
        xp[:, k] += h
        xm[:, k] -= h
        rp, rm = fwd(proc, xp), fwd(proc, xm)
        good = ok & rp["ok"] & rm["ok"]
        fd = (rp["p"] - rm["p"]) / (2 * h)
        sc = np.abs(fd[good]).max() + 1
        assert np.allclose(d["p_tan"][good][..., k], fd[good], atol=1e-5 * sc, rtol=1e-5), k
        fdf = (rp["factor"] - rm["factor"]) / (2 * h)
        assert np.allclose(d["factor_tan"][good, k], fdf[good], rtol=1e-5, atol=1e-9), k
    for j, arr, sel in ((proc.n_x, "map_masses", mapped), (proc.n_x + 1, "map_widths", proc.map_ids == P.MAP_SCHANNEL)):
        pp, pm = proc.subset(range(proc.n_channels)), proc.subset(range(proc.n_channels))
        getattr(pp, arr)[sel] += h
        getattr(pm, arr)[sel] -= h
        rp, rm = fwd(pp, x), fwd(pm, x)
        good = ok & rp["ok"] & rm["ok"]
        fdf = (rp["factor"] - rm["factor"]) / (2 * h)
        assert np.allclose(d["factor_tan"][good, j], fdf[good], rtol=1e-5, atol=1e-9), arr
        fd = (rp["p"] - rm["p"]) / (2 * h)
        assert np.allclose(d["p_tan"][good][..., j], fd[good], rtol=1e-5, atol=1e-5), arr


def test_dual_numbers_match_finite_differences():
    """The forward-mode reference of the adjoint: d p / d x and d factor / d(x, M, Gamma)."""
    _dual_against_differences(W.process_c1(), 200, 4)


@pytest.mark.parametrize("name,beams", [("c1", True), ("friends4", False), ("friends4", True)])
def test_dual_numbers_friends_and_event_beams(name, beams):
    """The dual numbers through the friend axis and with per-event beam energies (each event's
    map constants formed from the dual masses at its own sqrt(s))."""
    proc = W.process_c1() if name == "c1" else W.process_friends(4)
    _dual_against_differences(proc, 200, 9, beams=beams)


def test_per_event_beams_and_lab_boost():
    """Per-event sqrt(s) reduces to the fixed path for E1 = E2; for E1 != E2 the lab boost
    (phs.cu:1286-1301) keeps every p^2 and gives total momentum (E1 + E2, 0, 0, E1 - E2)."""
    proc = W.process_c1()
    rng = np.random.default_rng(8)
    n = 2000
    ch = rng.integers(0, 3, n).astype(np.int32)
    x = rng.random((n, proc.n_x))
    e = np.full(n, 250.0)
    a = oracle.forward(proc, x, ch)
    b = oracle.forward(proc, x, ch, E1=e, E2=e)
    assert np.array_equal(a["p"], b["p"]) and np.array_equal(a["factor"], b["factor"])
    E1 = 250.0 + 100 * rng.random(n)
    E2 = 250.0 - 100 * rng.random(n)
    c = oracle.forward(proc, x, ch, E1=E1, E2=E2)
    ok = c["ok"]
    tot = c["p"][ok].sum(axis=1)
    assert np.allclose(tot[:, 0], (E1 + E2)[ok], rtol=1e-11)
    assert np.allclose(tot[:, 3], (E1 - E2)[ok], atol=1e-8 * 500)
    assert np.allclose(tot[:, 1:3], 0, atol=1e-8 * 500)
    m2 = c["p"][ok][..., 0] ** 2 - (c["p"][ok][..., 1:] ** 2).sum(-1)
    assert np.allclose(m2, 0, atol=1e-7 * 500 ** 2)
