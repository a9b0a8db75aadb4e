"""ORACLE — test infrastructure only.

ctypes bindings of the CPU restatement (``liboracle.so``, built from phs_oracle.cpp) and,
where it was built, of the reference's own phs_cpu.cpp (``_ref/libphs_ref.so``).  Only
tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may import this
package, and only to check or to time the CPU baseline; the product path never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")   # ORACLE_LIB: the sanitizer build
REF_LIB = os.path.join(HERE, "_ref", "libphs_ref.so")
REF_SRC = "/root/reference"

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_long
_D = ctypes.c_double


def build(force: bool = False) -> None:
    """Compile liboracle.so, and _ref/libphs_ref.so where the reference tree is present."""
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-B" if force else "-s", "liboracle.so"], cwd=HERE, check=True)
    if os.path.isdir(REF_SRC) and (force or not os.path.exists(REF_LIB)):
        subprocess.run(["make", "-B" if force else "-s", "ref"], cwd=HERE, check=True)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        _lib.oracle_forward.restype = _I
        _lib.oracle_forward_friends.restype = _I
        _lib.oracle_forward_dual.restype = _I
        _lib.oracle_constants.restype = _I
        _lib.oracle_mass_sums.restype = _I
        _lib.oracle_map_msq.restype = None
        _lib.oracle_map_ct.restype = None
        _lib.oracle_map_msq.argtypes = [_I, _D, _D, _D, _D, _D, _D, _P, _P, _P]
        _lib.oracle_map_ct.argtypes = [_I, _D, _P, _P, _P, _P]
        _lib.oracle_set_perturbation.restype = None
        _lib.oracle_set_perturbation.argtypes = [ctypes.c_uint64]
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_LIB)
        _ref.ref_forward.restype = _I
        _ref.ref_time_forward.restype = _D
        _ref.ref_time_forward_mt.restype = _D
    return _ref


def _ptr(a):
    return a.ctypes.data_as(_P) if a is not None else None


def _tree_args(proc):
    d1, d2, hc, mid, mm, mw, flv = proc.ctypes_args()
    return [d1, d2, hc, mid, mm, mw, flv]


def forward(proc, x, channels, E1=None, E2=None, want_msq=False, perturb=0):
    """Oracle forward pass: returns dict(p, factor, volume, ok, scale, cond[, msq]); ``scale`` is
    the componentwise forward-error scale |L1|...|L1||v| of every momentum component, ``cond``
    the event's condition number (lambda near a threshold, sin theta of peaked maps, the
    friend axis).  Branches with a friend (``proc.friends``) take their polar axis from the
    incoming beam.  ``perturb`` != 0 moves every elementary-function result by up to one ulp
    along that seed's per-event streams (see ``sensitivity``)."""
    x = np.ascontiguousarray(x, np.float64)
    ch = np.ascontiguousarray(channels, np.int32)
    n = ch.shape[0]
    assert x.shape == (n, proc.n_x)
    p = np.zeros((n, proc.n_out, 4))
    f = np.zeros(n)
    v = np.zeros(n)
    ok = np.zeros(n, np.uint8)
    msq = np.zeros((n, proc.n_prt_out)) if want_msq else None
    scale = np.zeros((n, proc.n_out, 4))
    cond = np.zeros(n)
    t = _tree_args(proc)
    e1 = None if E1 is None else np.ascontiguousarray(E1, np.float64)
    e2 = None if E2 is None else np.ascontiguousarray(E2, np.float64)
    fr = None
    if proc.friends is not None and np.any(proc.friends):
        fr = np.ascontiguousarray(proc.friends, np.int32)
    lib().oracle_set_perturbation(ctypes.c_uint64(perturb))
    try:
        rc = lib().oracle_forward_friends(_I(proc.n_in), _I(proc.n_out), _I(proc.n_channels),
                                          *[_ptr(a) for a in t[:3]], _ptr(fr), *[_ptr(a) for a in t[3:]],
                                          _D(proc.sqrts), _ptr(e1), _ptr(e2), _L(n), _ptr(ch),
                                          _I(proc.n_x), _ptr(x), _ptr(p), _ptr(f), _ptr(v), _ptr(ok),
                                          _ptr(msq), _ptr(scale), _ptr(cond))
    finally:
        lib().oracle_set_perturbation(ctypes.c_uint64(0))
    if rc:
        raise ValueError(f"oracle_forward failed ({rc})")
    out = dict(p=p, factor=f, volume=v, ok=ok.astype(bool), scale=scale, cond=cond)
    if want_msq:
        out["msq"] = msq
    return out


def sensitivity(proc, x, channels, E1=None, E2=None, runs=6, base=None, want_msq=False):
    """Measured one-ulp sensitivity of every event: the oracle is run ``runs`` times with its
    elementary-function results (sqrt, exp, log, sin, cos, tan, atan) moved by -1, 0 or +1 ulp
    along independent seeded streams, and compared with the unperturbed run ``base`` (computed
    if not given).  Returns dict(p [n][n_out][4], factor [n]: the largest absolute change of
    each output over the runs; ok_flip [n]: the ok flag changed in some run, i.e. the event sits
    within rounding of a kinematic cut, and its outputs cannot be compared across the runs:
    their sensitivity is reported as infinite; with ``want_msq`` also msq [n][n_prt_out])."""
    if base is None or (want_msq and "msq" not in base):
        base = forward(proc, x, channels, E1=E1, E2=E2, want_msq=want_msq)
    sp = np.zeros_like(base["p"])
    sf = np.zeros_like(base["factor"])
    sm = np.zeros_like(base["msq"]) if want_msq else None
    flip = np.zeros(base["ok"].shape, bool)
    for k in range(runs):
        o = forward(proc, x, channels, E1=E1, E2=E2, want_msq=want_msq, perturb=0x5EED0000 + 7919 * (k + 1))
        flip |= o["ok"] != base["ok"]
        both = (o["ok"] & base["ok"])
        sp = np.maximum(sp, np.where(both[:, None, None], np.abs(o["p"] - base["p"]), 0.0))
        sf = np.maximum(sf, np.where(both, np.abs(o["factor"] - base["factor"]), 0.0))
        if want_msq:
            sm = np.maximum(sm, np.where(both[:, None], np.abs(o["msq"] - base["msq"]), 0.0))
    sp[flip] = np.inf
    sf[flip] = np.inf
    out = dict(p=sp, factor=sf, ok_flip=flip)
    if want_msq:
        sm[flip] = np.inf
        out["msq"] = sm
    return out


def inverse(proc, p, E1=None, E2=None):
    """Inverse mapping of events with momenta p [n][n_out][4] into every channel: returns
    (factors [n][n_channels], x_out [n_channels][n][n_x])."""
    p = np.ascontiguousarray(p, np.float64)
    n = p.shape[0]
    assert p.shape == (n, proc.n_out, 4)
    f = np.zeros((n, proc.n_channels))
    xo = np.zeros((proc.n_channels, n, proc.n_x))
    t = _tree_args(proc)
    fr = None
    if proc.friends is not None and np.any(proc.friends):
        fr = np.ascontiguousarray(proc.friends, np.int32)
    e1 = None if E1 is None else np.ascontiguousarray(E1, np.float64)
    e2 = None if E2 is None else np.ascontiguousarray(E2, np.float64)
    lib().oracle_inverse.restype = _I
    rc = lib().oracle_inverse(_I(proc.n_in), _I(proc.n_out), _I(proc.n_channels), *[_ptr(a) for a in t[:3]],
                              _ptr(fr), *[_ptr(a) for a in t[3:]], _D(proc.sqrts), _ptr(e1), _ptr(e2), _L(n),
                              _I(proc.n_x), _ptr(p), _ptr(f), _ptr(xo))
    if rc:
        raise ValueError(f"oracle_inverse failed ({rc})")
    return f, xo


def forward_dual(proc, x, channels, x_tan=None, m_tan=None, w_tan=None, K=1, E1=None, E2=None):
    """Dual-number forward: tangents of p and factor along K seeded directions (friend trees
    and per-event beam energies E1, E2 included)."""
    x = np.ascontiguousarray(x, np.float64)
    ch = np.ascontiguousarray(channels, np.int32)
    n = ch.shape[0]
    p = np.zeros((n, proc.n_out, 4))
    f = np.zeros(n)
    ok = np.zeros(n, np.uint8)
    pt = np.zeros((n, proc.n_out, 4, K))
    ft = np.zeros((n, K))
    cvt = (lambda a: None if a is None else np.ascontiguousarray(a, np.float64))
    xt, mt, wt = cvt(x_tan), cvt(m_tan), cvt(w_tan)
    t = _tree_args(proc)
    fr = None if proc.friends is None else np.ascontiguousarray(proc.friends, np.int32)
    e1 = None if E1 is None else np.ascontiguousarray(E1, np.float64)
    e2 = None if E2 is None else np.ascontiguousarray(E2, np.float64)
    rc = lib().oracle_forward_dual(_I(proc.n_in), _I(proc.n_out), _I(proc.n_channels),
                                   *[_ptr(a) for a in t[:3]], _ptr(fr), *[_ptr(a) for a in t[3:]],
                                   _D(proc.sqrts), _ptr(e1), _ptr(e2), _L(n), _ptr(ch),
                                   _I(proc.n_x), _ptr(x), _I(K), _ptr(xt), _ptr(mt), _ptr(wt),
                                   _ptr(p), _ptr(f), _ptr(ok), _ptr(pt), _ptr(ft))
    if rc:
        raise ValueError(f"oracle_forward_dual failed ({rc})")
    return dict(p=p, factor=f, ok=ok.astype(bool), p_tan=pt, factor_tan=ft)


def constants(proc, sqrts=None):
    n = proc.n_channels * proc.n_prt_out * 3
    a = np.zeros(n)
    b = np.zeros(n)
    t = _tree_args(proc)
    rc = lib().oracle_constants(_I(proc.n_in), _I(proc.n_out), _I(proc.n_channels),
                                *[_ptr(q) for q in t], _D(proc.sqrts if sqrts is None else sqrts),
                                _ptr(a), _ptr(b))
    if rc:
        raise ValueError("oracle_constants failed")
    shape = (proc.n_channels, proc.n_prt_out, 3)
    return a.reshape(shape), b.reshape(shape)


def mass_sums(proc):
    out = np.zeros(proc.n_channels * proc.n_prt_out)
    d1, d2, hc, mid, mm, mw, flv = proc.ctypes_args()
    rc = lib().oracle_mass_sums(_I(proc.n_in), _I(proc.n_out), _I(proc.n_channels), _ptr(d1),
                                _ptr(d2), _ptr(hc), _ptr(mid), _ptr(flv), _ptr(out))
    if rc:
        raise ValueError("oracle_mass_sums failed")
    return out.reshape(proc.n_channels, proc.n_prt_out)


def map_msq(map_id, x, s, m, w, msq_min, msq_max, a):
    a = np.ascontiguousarray(a, np.float64)
    msq = ctypes.c_double()
    f = ctypes.c_double()
    lib().oracle_map_msq(map_id, x, s, m, w, msq_min, msq_max, _ptr(a), ctypes.byref(msq),
                         ctypes.byref(f))
    return msq.value, f.value


def map_ct(map_id, x, b):
    b = np.ascontiguousarray(b, np.float64)
    ct, st, f = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().oracle_map_ct(map_id, x, _ptr(b), ctypes.byref(ct), ctypes.byref(st), ctypes.byref(f))
    return ct.value, st.value, f.value


def ref_forward(proc, x, channels, want_msq=False):
    """The compiled reference phs_cpu.cpp on the same inputs (fixed sqrt(s))."""
    x = np.array(x, np.float64, order="C")  # the reference takes a non-const pointer
    ch = np.ascontiguousarray(channels, np.int32)
    n = ch.shape[0]
    p = np.zeros((n, proc.n_out, 4))
    f = np.zeros(n)
    v = np.zeros(n)
    ok = np.zeros(n, np.uint8)
    msq = np.zeros((n, proc.n_prt_out)) if want_msq else None
    t = _tree_args(proc)
    rc = ref().ref_forward(_I(proc.n_in), _I(proc.n_out), _I(proc.n_channels),
                           *[_ptr(a) for a in t], _D(proc.sqrts), _L(n), _ptr(ch), _I(proc.n_x),
                           _ptr(x), _ptr(p), _ptr(f), _ptr(v), _ptr(ok), _ptr(msq))
    if rc:
        raise ValueError(f"ref_forward failed ({rc})")
    out = dict(p=p, factor=f, volume=v, ok=ok.astype(bool))
    if want_msq:
        out["msq"] = msq
    return out


def ref_time(proc, x, channels):
    """Seconds the reference's single-thread generation loop takes, and its ok count."""
    x = np.array(x, np.float64, order="C")
    ch = np.ascontiguousarray(channels, np.int32)
    n_ok = ctypes.c_long()
    t = _tree_args(proc)
    sec = ref().ref_time_forward(_I(proc.n_in), _I(proc.n_out), _I(proc.n_channels),
                                 *[_ptr(a) for a in t], _D(proc.sqrts), _L(ch.shape[0]), _ptr(ch),
                                 _I(proc.n_x), _ptr(x), ctypes.byref(n_ok))
    return sec, n_ok.value


def ref_time_threads(proc, x, channels, threads):
    """Wall seconds of the reference's generation loop split over `threads` host threads
    (contiguous event ranges), and its ok count."""
    x = np.array(x, np.float64, order="C")
    ch = np.ascontiguousarray(channels, np.int32)
    n_ok = ctypes.c_long()
    t = _tree_args(proc)
    sec = ref().ref_time_forward_mt(_I(proc.n_in), _I(proc.n_out), _I(proc.n_channels),
                                    *[_ptr(a) for a in t], _D(proc.sqrts), _L(ch.shape[0]), _ptr(ch),
                                    _I(proc.n_x), _ptr(x), _I(threads), ctypes.byref(n_ok))
    if sec < 0:
        raise RuntimeError("reference timing failed")
    return sec, n_ok.value


# ---------------------------------------------------------------------------------------------
# Adaptive grids of the randoms (rng.cpp:18-93), numpy restatement of the device version.

GRID_BINS = 10   # rng.cpp:5


def philox4x32_10(ctr, key):
    """Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC 2011): ctr [n, 4] uint32, key (k0, k1)."""
    c = [np.asarray(ctr[:, i], np.uint64) for i in range(4)]
    k0, k1 = np.uint64(key[0]), np.uint64(key[1])
    M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
    W0, W1 = np.uint64(0x9E3779B9), np.uint64(0xBB67AE85)
    mask = np.uint64(0xFFFFFFFF)
    for _ in range(10):
        p0, p1 = M0 * c[0], M1 * c[2]
        hi0, lo0 = p0 >> np.uint64(32), p0 & mask
        hi1, lo1 = p1 >> np.uint64(32), p1 & mask
        c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
        k0, k1 = (k0 + W0) & mask, (k1 + W1) & mask
    return np.stack(c, axis=1).astype(np.uint32)


def grid_cdf(w):
    """Cumulative bin probabilities [..., 11] of weights [..., 10], F[10] = 1."""
    w = np.asarray(w, np.float64)
    s = np.zeros(w.shape[:-1])
    for b in range(GRID_BINS):
        s = s + w[..., b]
    F = np.zeros(w.shape[:-1] + (GRID_BINS + 1,))
    for b in range(GRID_BINS):
        F[..., b + 1] = F[..., b] + w[..., b] / s
    F[..., GRID_BINS] = 1.0
    return F


def grid_uniforms(seed, segment, n_events, n_x):
    """The uniform numbers u [n_events, n_x] of one segment: random k of event e from the
    Philox block (e, segment, k // 2), two 32-bit words per double."""
    n_blk = (n_x + 1) // 2
    e = np.repeat(np.arange(n_events, dtype=np.uint64), n_blk)
    j = np.tile(np.arange(n_blk, dtype=np.uint64), n_events)
    ctr = np.stack([e & np.uint64(0xFFFFFFFF), e >> np.uint64(32), np.full_like(e, segment), j], axis=1)
    out = philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32)).astype(np.uint64).reshape(n_events, n_blk, 4)
    u = np.empty((n_events, 2 * n_blk))
    for h in range(2):
        bits = ((out[:, :, 2 * h] << np.uint64(32)) | out[:, :, 2 * h + 1]) >> np.uint64(11)
        u[:, h::2] = bits.astype(np.float64) * 2.0 ** -53
    return u[:, :n_x]


def grid_sample(F, u):
    """x from u through the inverse distribution function of the piecewise-constant density
    with cumulative bin probabilities F [n_x, 11] (rng.cpp:74-79: 10 equal bins on [0, 1]):
    bin b = the first with F[b + 1] > u, x = b / 10 + (u - F[b]) / rho_b."""
    u = np.asarray(u, np.float64)
    x = np.empty_like(u)
    for k in range(u.shape[1]):
        Fk = F[k]
        b = np.zeros(u.shape[0], np.int64)
        for i in range(GRID_BINS - 1):
            b += (b == i) & ~(Fk[i + 1] > u[:, k])
        rho = (Fk[b + 1] - Fk[b]) * GRID_BINS
        x[:, k] = b / GRID_BINS + (u[:, k] - Fk[b]) / rho
    return x


def grid_update(w, x, ok, channels):
    """rng.cpp:33-67: per channel with accepted events, weight = bin count / accepted count;
    the bin of x is (int)(x * 10).  Channels without accepted events keep w."""
    w = np.array(w, np.float64)
    n_ch, n_dim, _ = w.shape
    for c in range(n_ch):
        sel = ok & (channels == c)
        n_ok = int(sel.sum())
        if n_ok == 0:
            continue
        b = np.clip((x[sel] * GRID_BINS).astype(np.int64), 0, GRID_BINS - 1)
        for d in range(n_dim):
            w[c, d] = np.bincount(b[:, d], minlength=GRID_BINS) / n_ok
    return w
