"""The kernels' bounded-range sin/cos/tan/exp/log (gpu_phs_amd/csrc/phs_math.h), compiled for the host,
against 300-bit arithmetic (mpmath).  The kernels use the same code, so the error bounds
checked here are the ones the GPU parity tests rely on."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "gpu_phs_amd", "csrc")
mp = pytest.importorskip("mpmath")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("math") / "libmath_host.so")
    # -mfma: the fused multiply-adds become instructions, as on the GPU (without it they are
    # libm calls with the same, correctly rounded, result)
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-I", CSRC,
                           os.path.join(HERE, "math_host.cpp"), "-o", out])
    L = ctypes.CDLL(out)
    P = ctypes.POINTER(ctypes.c_double)
    L.host_sincos_turns.argtypes = [P, ctypes.c_long, P, P]
    L.host_tan_radians.argtypes = [P, ctypes.c_long, P]
    L.host_exp.argtypes = [P, ctypes.c_long, P]
    L.host_log.argtypes = [P, ctypes.c_long, P]
    L.host_atan2_turns.argtypes = [P, P, ctypes.c_long, P]
    L.host_atan.argtypes = [P, ctypes.c_long, P]
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ulps(got, exact):
    """|got - exact| in units of the last place of the exact value (exact as mpf)."""
    out = np.empty(len(got))
    for i, (g, e) in enumerate(zip(got, exact)):
        ef = float(e)
        if ef == 0:
            out[i] = 0.0 if g == 0 else np.inf
            continue
        out[i] = float(abs(mp.mpf(g) - e)) / np.spacing(abs(ef))
    return out


def _turn_points():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.random(3000),
                        # quadrant edges and their neighbours, where reduction errors would show
                        np.array([k / 8 for k in range(9)]),
                        np.nextafter(np.array([k / 8 for k in range(1, 9)]), 0),
                        np.nextafter(np.array([k / 8 for k in range(0, 8)]), 1),
                        np.array([1e-300, 5e-324, 1e-17, 0.5 - 1e-17, 1 - 2 ** -53]),
                        -rng.random(200), 1 + 3 * rng.random(200)])
    return x


def test_sincos_turns_within_two_ulps(lib):
    mp.mp.prec = 300
    x = _turn_points()
    s = np.empty_like(x)
    c = np.empty_like(x)
    lib.host_sincos_turns(_p(x), len(x), _p(s), _p(c))
    # sin(2 pi x) and cos(2 pi x) are exactly 0 or +-1 where 4x is an integer
    quarter = {0: (0, 1), 1: (1, 0), 2: (0, -1), 3: (-1, 0)}
    es = [mp.mpf(quarter[int(4 * v) % 4][0]) if 4 * v == int(4 * v) else mp.sin(2 * mp.pi * mp.mpf(v)) for v in x]
    ec = [mp.mpf(quarter[int(4 * v) % 4][1]) if 4 * v == int(4 * v) else mp.cos(2 * mp.pi * mp.mpf(v)) for v in x]
    us, uc = _ulps(s, es), _ulps(c, ec)
    print("max ulps sin", us.max(), x[us.argmax()], "cos", uc.max(), x[uc.argmax()])
    # exact zeros at the quadrant edges
    for k in range(5):
        xs = np.array([k / 4])
        lib.host_sincos_turns(_p(xs), 1, _p(s[:1]), _p(c[:1]))
        assert (s[0], c[0]) == [(0.0, 1.0), (1.0, 0.0), (0.0, -1.0), (-1.0, 0.0), (0.0, 1.0)][k]
    assert us.max() <= 2.0 and uc.max() <= 2.0


def test_tan_radians_within_three_ulps(lib):
    mp.mp.prec = 300
    rng = np.random.default_rng(2)
    half_pi = np.pi / 2
    z = np.concatenate([(2 * rng.random(3000) - 1) * half_pi,
                        (2 * rng.random(300) - 1) * 1e-3,
                        np.array([half_pi, -half_pi, np.nextafter(half_pi, 0), np.pi / 4, -np.pi / 4,
                                  np.nextafter(np.pi / 4, 1), 1e-300, 0.0, -0.0, 1.2, -1.5]),
                        (2 * rng.random(300) - 1) * 20])
    t = np.empty_like(z)
    lib.host_tan_radians(_p(z), len(z), _p(t))
    u = _ulps(t, [mp.tan(mp.mpf(v)) for v in z])
    print("max ulps tan", u.max(), z[u.argmax()])
    assert u.max() <= 3.0


def test_exp_within_two_ulps(lib):
    mp.mp.prec = 300
    rng = np.random.default_rng(3)
    x = np.concatenate([(2 * rng.random(2000) - 1) * 40, (2 * rng.random(500) - 1) * 700,
                        (2 * rng.random(300) - 1) * 1e-6,
                        np.array([0.0, -0.0, 1e-300, 709.7, -744.0, -708.5, 0.5 * np.log(2), -0.5 * np.log(2)])])
    y = np.empty_like(x)
    lib.host_exp(_p(x), len(x), _p(y))
    u = _ulps(y, [mp.exp(mp.mpf(v)) for v in x])
    print("max ulps exp", u.max(), x[u.argmax()])
    assert u.max() <= 2.0
    # the ends of the range and special values
    sp = np.array([710.0, 1e308, -746.0, -1e308, np.inf, -np.inf, np.nan])
    ys = np.empty_like(sp)
    lib.host_exp(_p(sp), len(sp), _p(ys))
    assert ys[0] == np.inf and ys[1] == np.inf and ys[2] == 0 and ys[3] == 0
    assert ys[4] == np.inf and ys[5] == 0 and np.isnan(ys[6])


def test_log_within_two_ulps(lib):
    mp.mp.prec = 300
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.random(1500) * 4, 10.0 ** (rng.random(1000) * 600 - 300),
                        1 + (2 * rng.random(300) - 1) * 1e-8,
                        np.array([1.0, 2.0, 0.5, np.sqrt(0.5), np.nextafter(np.sqrt(0.5), 1), 5e-324, 1.7e308])])
    y = np.empty_like(x)
    lib.host_log(_p(x), len(x), _p(y))
    u = _ulps(y, [mp.log(mp.mpf(v)) for v in x])
    print("max ulps log", u.max(), x[u.argmax()])
    assert u.max() <= 2.0
    sp = np.array([0.0, -1.0, np.inf, np.nan])
    ys = np.empty_like(sp)
    lib.host_log(_p(sp), len(sp), _p(ys))
    assert ys[0] == -np.inf and np.isnan(ys[1]) and ys[2] == np.inf and np.isnan(ys[3])


def test_atan2_turns_within_three_ulps(lib):
    """The azimuth in turns, atan2(y, x) / 2 pi (+ 1 where negative), against 300-bit
    arithmetic, over the whole circle: random directions and lengths, the octant and half-plane
    edges, the edges of the reduction's intervals (tan of (2k + 1) pi / 16... a = (2k + 1) / 8),
    and tiny and huge components.  (The host build forms its quotients as a reciprocal and a
    product, as the device does.)"""
    mp.mp.prec = 300
    rng = np.random.default_rng(5)
    ang = rng.random(3000) * 2 * np.pi
    r = 10.0 ** (rng.random(3000) * 8 - 4)
    y, x = r * np.sin(ang), r * np.cos(ang)
    edges = []
    for a in [0.0, 1 / 8, 3 / 8, 5 / 8, 7 / 8, 1.0, 0.25, 0.5, 0.75]:
        for aa in (a, np.nextafter(a, 0), np.nextafter(a, 2)):
            for sx in (1, -1):
                for sy in (1, -1):
                    edges += [(sy * aa, sx * 1.0), (sy * 1.0, sx * aa)]
    edges += [(1e-300, 1.0), (-1e-300, 1.0), (1e-300, -1.0), (-1e-300, -1.0), (3.0, 1e-310), (2.0, -5e-324),
              (1e200, 3e199), (-7e-200, 2e-200)]
    ey, ex = np.array(edges).T
    y, x = np.concatenate([y, ey]), np.concatenate([x, ex])
    t = np.empty_like(x)
    lib.host_atan2_turns(_p(y), _p(x), len(x), _p(t))
    exact = []
    for yy, xx in zip(y, x):
        v = mp.atan2(mp.mpf(yy), mp.mpf(xx)) / (2 * mp.pi)
        exact.append(v + 1 if v < 0 else v)
    u = _ulps(t, exact)
    print("max ulps atan2_turns", u.max(), y[u.argmax()], x[u.argmax()])
    assert u.max() <= 3.0
    # signed zeros as atan2 places them, and NaN
    sp = [(0.0, 1.0, 0.0), (0.0, -1.0, 0.5), (-0.0, -1.0, 0.5), (0.0, 0.0, 0.0), (0.0, -0.0, 0.5),
          (1.0, 0.0, 0.25), (-1.0, 0.0, 0.75), (1.0, 1.0, 0.125), (-1.0, -1.0, 0.625)]
    ys, xs, want = (np.array(v) for v in zip(*sp))
    out = np.empty_like(xs)
    lib.host_atan2_turns(_p(ys), _p(xs), len(xs), _p(out))
    assert np.array_equal(out, want)
    nan = np.array([np.nan, 1.0]), np.array([1.0, np.nan])
    out = np.empty(2)
    lib.host_atan2_turns(_p(nan[0]), _p(nan[1]), 2, _p(out))
    assert np.isnan(out).all()


def test_atan_radians_within_0_6_ulps(lib):
    """atan in radians (the Breit-Wigner map's constants at each event's sqrt(s)) against
    300-bit arithmetic: random arguments over 2e-12 .. 2e12 of both signs, the edges of the
    reduction's intervals (a = (2k + 1) / 8 and their neighbours, 1 and its neighbours, and the
    same as reciprocals), tiny, huge and infinite arguments; signed zeros and NaN.  A dropped
    series term, a wrong table entry or a wrong branch of the reduction shows here as hundreds of
    ulps."""
    mp.mp.prec = 300
    rng = np.random.default_rng(9)
    x = 10.0 ** (rng.random(4000) * 24 - 12) * rng.choice([-1.0, 1.0], 4000)
    edges = []
    for a in [1 / 8, 3 / 8, 5 / 8, 7 / 8, 0.25, 0.5, 0.75, 1.0]:
        for aa in (a, np.nextafter(a, 0), np.nextafter(a, 2)):
            edges += [aa, -aa, 1 / aa, -1 / aa]
    edges += [1e-300, -5e-324, 1e-17, 3e15, 1e300, -1.7e308]
    x = np.concatenate([x, np.array(edges)])
    y = np.empty_like(x)
    lib.host_atan(_p(x), len(x), _p(y))
    u = _ulps(y, [mp.atan(mp.mpf(v)) for v in x])
    print("max ulps atan_radians", u.max(), x[u.argmax()], "correctly rounded", np.mean(u <= 0.5))
    assert u.max() <= 0.6
    sp = np.array([0.0, -0.0, np.inf, -np.inf, np.nan])
    out = np.empty_like(sp)
    lib.host_atan(_p(sp), len(sp), _p(out))
    assert out[0] == 0 and not np.signbit(out[0]) and out[1] == 0 and np.signbit(out[1])
    assert out[2] == np.pi / 2 and out[3] == -np.pi / 2 and np.isnan(out[4])
