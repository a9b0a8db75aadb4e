"""The Whizard-facing C ABI on the GPU: the reference's call sequence (whizard_gpu_phs.cpp)
with host arrays, checked against the oracle and against Whizard's own momenta."""
import ctypes
import os

import numpy as np
import pytest

import gpu_phs_amd as G
import oracle
from gpu_phs_amd import workloads as W
from parity import assert_parity, forward_check
from refio import read_ref
from test_gpu_inverse import assert_inverse_parity, inverse_stats, sensitivity

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _i(v):
    return ctypes.byref(ctypes.c_int(v))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def setup_module_from(proc, block=256):
    L = G.lib()
    nc = proc.n_channels
    L.c_whizard_set_particle_structure(_i(nc), _i(nc), _i(1), _i(proc.n_x), _i(proc.n_in), _i(proc.n_out))
    flv = np.ascontiguousarray(proc.flv_masses)
    L.c_whizard_set_flavors(_p(flv), _p(np.zeros_like(flv)))
    L.c_whizard_init_mappings()
    L.c_whizard_init_tree_structures()
    keep = []
    for c in range(nc):
        arrs = [np.ascontiguousarray(a[c]) for a in (proc.map_ids, proc.map_masses, proc.map_widths,
                                                     proc.d1, proc.d2, proc.has_children)]
        arrs.append(np.zeros(proc.n_prt, np.int32))
        keep.append(arrs)
        L.c_whizard_fill_mapping(_i(c + 1), _p(arrs[0]), _p(arrs[1]), _p(arrs[2]))
        L.c_whizard_fill_tree_structure(_i(c + 1), _p(arrs[3]), _p(arrs[4]), _p(arrs[5]), _p(arrs[6]))
    L.c_whizard_set_threads(_i(block), _i(block), _i(block))
    L.c_whizard_init_gpu_phs()
    assert L.gpu_phs_last_error() == 0, L.gpu_phs_last_error_message()
    return L


def gen(L, proc, x, counts, E1, E2):
    n = x.shape[0]
    limits = (ctypes.c_int * len(counts))(*np.cumsum(counts).tolist())
    L.c_whizard_init_channel_ids(_i(n), _i(len(counts)), limits)
    assert L.gpu_phs_last_error() == 0
    nc = proc.n_channels
    e1 = np.full(n, E1)
    e2 = np.full(n, E2)
    factors = np.full(n * nc, -1.0)
    volumes = np.zeros(n)
    oks = np.zeros(n, np.bool_)
    p = np.zeros((n, proc.n_out, 4))
    x_out = np.full(n * nc * proc.n_x, -1.0)
    flag = ctypes.c_bool(True)
    x = np.ascontiguousarray(x)
    L.c_whizard_gen_phs_from_x_gpu(_p(e1), _p(e2), ctypes.byref(flag), _i(n), _i(nc), _i(proc.n_x), _p(x),
                                   _p(factors), _p(volumes), _p(oks), _p(p), _p(x_out))
    assert L.gpu_phs_last_error() == 0, L.gpu_phs_last_error_message()
    return p, factors.reshape(n, nc), volumes, oks, x_out.reshape(nc, n, proc.n_x)


def test_whizard_abi_matches_oracle():
    proc = W.process_c1()
    L = setup_module_from(proc)
    counts = [0, 700, 1300]          # an empty channel first, as Whizard's batches allow
    n = sum(counts)
    rng = np.random.default_rng(21)
    x = rng.random((n, proc.n_x))
    p, factors, volumes, oks, x_out = gen(L, proc, x, counts, 250.0, 250.0)
    ch = np.repeat(np.arange(3), counts).astype(np.int32)
    own = factors[np.arange(n), ch]
    assert_parity(forward_check(proc, x, ch, dict(p=p, factor=own, volume=volumes, ok=oks)))
    # every other channel's factor and randoms: the inverse mapping of the event's momenta
    for c in range(3):
        sel = ch == c
        assert np.array_equal(x_out[c][sel], x[sel])
    p_ok = np.where(oks[:, None, None], p, 0.0)
    f_ref, x_ref = oracle.inverse(proc, p_ok)
    df, dx = inverse_stats(proc, factors, x_out, f_ref, x_ref, ch, oks)
    assert_inverse_parity(df, dx, *sensitivity(proc, p_ok, f_ref, x_ref, ch, oks))
    # the device pointer of the last call's momenta
    addr = np.zeros(1, np.float64)
    L.c_whizard_get_momentum_device_pointer(_p(addr))
    assert addr.view(np.uint64)[0] != 0
    L.c_whizard_finalize_gpu_phs()


def test_whizard_abi_interior_empty_channel():
    proc = W.process_c1()
    L = setup_module_from(proc, block=64)
    counts = [500, 0, 321]
    n = sum(counts)
    x = np.random.default_rng(5).random((n, proc.n_x))
    p, factors, volumes, oks, _ = gen(L, proc, x, counts, 250.0, 250.0)
    ch = np.repeat(np.arange(3), counts).astype(np.int32)
    assert_parity(forward_check(proc, x, ch, dict(p=p, factor=factors[np.arange(n), ch], volume=volumes, ok=oks)))
    L.c_whizard_finalize_gpu_phs()


def test_whizard_abi_asymmetric_beams_lab_boost():
    proc = W.process_c1()
    L = setup_module_from(proc)
    counts = [400, 400, 400]
    n = sum(counts)
    x = np.random.default_rng(6).random((n, proc.n_x))
    E1, E2 = 300.0, 200.0
    p, factors, volumes, oks, _ = gen(L, proc, x, counts, E1, E2)
    ch = np.repeat(np.arange(3), counts).astype(np.int32)
    assert_parity(forward_check(proc, x, ch, dict(p=p, factor=factors[np.arange(n), ch], volume=volumes, ok=oks),
                                E1=np.full(n, E1), E2=np.full(n, E2)))
    tot = p[oks].sum(axis=1)
    assert np.allclose(tot[:, 0], E1 + E2) and np.allclose(tot[:, 3], E1 - E2, atol=1e-9)
    L.c_whizard_finalize_gpu_phs()


@pytest.mark.parametrize("name", ["ttH_head.ref", "bwbw_head.ref"])
def test_reference_samples_on_gpu(name):
    """Whizard's own momenta (9-decimal randoms) from the reference's sample files, through the
    Whizard ABI on the GPU; trees with friends are left out here (the file carries no per-branch
    friend data; test_gpu_friend.py covers them with the assignment Whizard's factors fix)."""
    proc, ev = read_ref(os.path.join(GOLDEN, name))
    L = setup_module_from(proc)
    ch = ev["channel"]
    order = np.argsort(ch, kind="stable")
    counts = np.bincount(ch, minlength=proc.n_channels)
    p, factors, volumes, oks, _ = gen(L, proc, ev["x"][order], counts, proc.sqrts / 2, proc.sqrts / 2)
    sel = ev["hasfriends"][ch[order]] == 0
    assert oks[sel].all()
    pw = ev["p"][order][:, proc.n_in:, :]
    dev = np.abs(p[sel] - pw[sel]).max() / proc.sqrts
    assert dev < 1e-6, dev
    assert np.allclose(volumes[sel], ev["volume"][order][sel], rtol=1e-12)
    L.c_whizard_finalize_gpu_phs()


def test_whizard_abi_per_event_beams():
    """Beam energies that differ from event to event: each event its own sqrt(s)."""
    proc = W.process_c1()
    L = setup_module_from(proc)
    counts = [600, 0, 700]
    n = sum(counts)
    rng = np.random.default_rng(8)
    x = rng.random((n, proc.n_x))
    E1, E2 = 250.0 * rng.uniform(0.8, 1.3, n), 250.0 * rng.uniform(0.8, 1.3, n)
    ch = np.repeat(np.arange(3), counts).astype(np.int32)
    limits = (ctypes.c_int * 3)(*np.cumsum(counts).tolist())
    L.c_whizard_init_channel_ids(_i(n), _i(3), limits)
    nc = proc.n_channels
    factors, volumes = np.zeros(n * nc), np.zeros(n)
    oks, p = np.zeros(n, np.bool_), np.zeros((n, proc.n_out, 4))
    x_out = np.zeros(n * nc * proc.n_x)
    flag = ctypes.c_bool(True)
    L.c_whizard_gen_phs_from_x_gpu(_p(np.ascontiguousarray(E1)), _p(np.ascontiguousarray(E2)), ctypes.byref(flag),
                                   _i(n), _i(nc), _i(proc.n_x), _p(x), _p(factors), _p(volumes), _p(oks), _p(p),
                                   _p(x_out))
    assert L.gpu_phs_last_error() == 0, L.gpu_phs_last_error_message()
    factors = factors.reshape(n, nc)
    assert_parity(forward_check(proc, x, ch, dict(p=p, factor=factors[np.arange(n), ch], volume=volumes, ok=oks),
                                E1=E1, E2=E2))
    p_ok = np.where(oks[:, None, None], p, 0.0)
    f_ref, x_ref = oracle.inverse(proc, p_ok, E1=E1, E2=E2)
    df, dx = inverse_stats(proc, factors, x_out.reshape(nc, n, proc.n_x), f_ref, x_ref, ch, oks)
    assert_inverse_parity(df, dx, *sensitivity(proc, p_ok, f_ref, x_ref, ch, oks, E1=E1, E2=E2))
    L.c_whizard_finalize_gpu_phs()


def test_device_pointers_of_the_last_call():
    """transfer_to_whizard.h: the momenta (Whizard layout), beam energies and ok flags of the
    last generate call stay on the device."""
    proc = W.process_c1()
    L = setup_module_from(proc)
    counts = [300, 200, 100]
    n = sum(counts)
    x = np.random.default_rng(12).random((n, proc.n_x))
    E1, E2 = 260.0, 240.0
    p, factors, volumes, oks, _ = gen(L, proc, x, counts, E1, E2)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int

    def fetch(sym, shape, dtype):
        f = getattr(L, sym)
        f.restype = ctypes.c_void_p
        addr = f()
        assert addr
        out = np.zeros(shape, dtype)
        assert hip.hipMemcpy(out.ctypes.data, addr, out.nbytes, 2) == 0   # device to host
        return out

    assert np.array_equal(fetch("_Z27get_momentum_device_pointerv", p.shape, np.float64), p)
    beams = fetch("_Z30get_in_momentum_device_pointerv", (2, n), np.float64)
    assert (beams[0] == E1).all() and (beams[1] == E2).all()
    assert np.array_equal(fetch("_Z21get_ok_device_pointerv", n, np.bool_), oks)
    L.c_whizard_finalize_gpu_phs()


def test_whizard_abi_large_trees_at_block_256():
    """2->8 trees (two levels of LDS save banks) through the C ABI with set_threads(256): each
    tile class keeps a workgroup its LDS tables fit into."""
    proc = W.process_large(8)
    L = setup_module_from(proc, block=256)
    rng = np.random.default_rng(31)
    counts = [900] * proc.n_channels
    n = sum(counts)
    x = rng.random((n, proc.n_x))
    h = proc.sqrts / 2
    p, factors, volumes, oks, _ = gen(L, proc, x, counts, h, h)
    ch = np.repeat(np.arange(proc.n_channels), counts).astype(np.int32)
    r = forward_check(proc, x, ch, dict(p=p, factor=factors[np.arange(n), ch], volume=volumes, ok=oks))
    assert r["n_ok"] > 100
    assert_parity(r)
    L.c_whizard_finalize_gpu_phs()


def test_whizard_abi_repeated_calls_reuse_and_resplit():
    """Calls with the same batch split reuse the module's plan; a new split rebuilds it; every
    call matches the oracle, and a repeated call gives identical results."""
    proc = W.process_c1()
    L = setup_module_from(proc)
    rng = np.random.default_rng(41)
    outs = []
    for counts in ([400, 300, 500], [400, 300, 500], [10, 1000, 3], [400, 300, 500]):
        n = sum(counts)
        x = rng.random((n, proc.n_x)) if len(outs) != 1 else outs[0][0]
        p, factors, volumes, oks, x_out = gen(L, proc, x, counts, 250.0, 250.0)
        ch = np.repeat(np.arange(3), counts).astype(np.int32)
        assert_parity(forward_check(proc, x, ch, dict(p=p, factor=factors[np.arange(n), ch], volume=volumes, ok=oks)))
        outs.append((x, p, factors, oks, x_out))
    # the second call repeated the first one's randoms and split
    for a, b in zip(outs[0][1:], outs[1][1:]):
        assert np.array_equal(a, b, equal_nan=True)
    L.c_whizard_finalize_gpu_phs()


def test_whizard_abi_large_call_identical_to_plan():
    """A call whose arrays span several staging chunks (16 MiB, each copied on the host by
    several threads), with a partial last chunk: every output equals, bit for bit, the plan API's
    forward pass and inverse mapping of the same events on the device."""
    import torch
    proc = W.process_c1()
    L = setup_module_from(proc)
    counts = [100_003, 100_003, 100_003]
    n = sum(counts)
    rng = np.random.default_rng(33)
    x = rng.random((n, proc.n_x))
    print("large Whizard call", n, flush=True)
    p, factors, volumes, oks, x_out = gen(L, proc, x, counts, 250.0, 250.0)
    L.c_whizard_finalize_gpu_phs()
    ph = G.ProcessHandle(proc)
    xt = torch.from_numpy(x).cuda()
    bufs, segs, start = [], [], 0
    for c, k in enumerate(counts):
        b = G.EventBuffers.allocate(xt[start:start + k], proc.n_out)
        bufs.append(b)
        segs.append((ph, c, b))
        start += k
    plan = G.Plan(segs)
    plan.run()
    f_plan, x_plan = plan.inverse()
    torch.cuda.synchronize()
    cat = lambda k: torch.cat([getattr(b, k) for b in bufs]).cpu().numpy()
    assert np.array_equal(oks, cat("ok").astype(bool))
    assert np.array_equal(p, cat("p"), equal_nan=True)
    assert np.array_equal(volumes, cat("volume"))
    assert np.array_equal(factors, f_plan.cpu().numpy(), equal_nan=True)
    assert np.array_equal(x_out, x_plan.cpu().numpy(), equal_nan=True)
