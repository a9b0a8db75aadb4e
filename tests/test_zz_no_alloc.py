"""The allocation-free entry point gpu_phs_generate_noalloc, called through the C ABI: two
calls in a row on one GPU (the second smaller, so stale tables from the first would show) give
bit for bit what a one-segment plan (gpu_phs_plan_create + gpu_phs_plan_run) gives on the same
randoms, for a shape-specialised 2->4 tree, a shape-specialised 2->6 tree and a tree with a
friend (generic kernel); a call beyond the array's capacity is rejected before any launch."""
import ctypes

import numpy as np
import pytest

import gpu_phs_amd as G
from gpu_phs_amd import workloads as W

pytestmark = pytest.mark.gpu

_P = ctypes.c_void_p


def _lib():
    L = G.lib()
    f = L.gpu_phs_generate_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, _P, _P, _P, _P, _P, _P]
    return L


def _noalloc(L, ph, channel, x, stream):
    import torch
    n = x.shape[0]
    p = torch.full((n, ph.n_out, 4), -7.0, dtype=torch.float64, device=x.device)
    f = torch.full((n,), -7.0, dtype=torch.float64, device=x.device)
    v = torch.full((n,), -7.0, dtype=torch.float64, device=x.device)
    ok = torch.full((n,), 7, dtype=torch.uint8, device=x.device)
    rc = L.gpu_phs_generate_noalloc(ph.handle, channel, n, x.data_ptr(), p.data_ptr(), f.data_ptr(), v.data_ptr(),
                                    ok.data_ptr(), stream.cuda_stream)
    assert rc == 0, L.gpu_phs_last_error_message().decode()
    return p, f, v, ok


def _plan(ph, channel, x):
    buf = G.EventBuffers.allocate(x, ph.n_out)
    plan = G.Plan([(ph, channel, buf)])
    plan.run()
    return buf, plan


@pytest.mark.parametrize("name", ["c1", "2to6", "friends4"])
def test_noalloc_matches_plan_twice(name):
    import torch
    assert torch.cuda.is_available()
    L = _lib()
    proc = {"c1": W.process_c1, "2to6": W.process_2to6, "friends4": lambda: W.process_friends(4)}[name]()
    ph = G.ProcessHandle(proc)
    stream = torch.cuda.current_stream()
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1234)
    for n, channel in ((300_007, 1), (5_003, 0)):
        print("noalloc", name, n, channel, flush=True)
        x = torch.rand((n, proc.n_x), dtype=torch.float64, device="cuda", generator=gen)
        p, f, v, ok = _noalloc(L, ph, channel, x, stream)
        buf, plan = _plan(ph, channel, x)
        torch.cuda.synchronize()
        assert torch.equal(ok, buf.ok)
        assert (ok == 0).sum() + (ok == 1).sum() == n          # every event written
        m = ok.bool()
        assert int(m.sum()) > 0
        assert torch.equal(p[m], buf.p[m]) and torch.equal(f[m], buf.factor[m]) and torch.equal(v, buf.volume)
        assert torch.equal(torch.isnan(p), torch.isnan(buf.p))
        del plan


def test_noalloc_rejects_calls_beyond_capacity():
    import torch
    L = _lib()
    ph = G.ProcessHandle(W.process_2to6())
    x = torch.zeros((8, ph.n_x), dtype=torch.float64, device="cuda")
    big = 16_515_072 + 1       # one event past 64,512 tiles of 256 events
    rc = L.gpu_phs_generate_noalloc(ph.handle, 0, big, x.data_ptr(), x.data_ptr(), x.data_ptr(), x.data_ptr(),
                                    x.data_ptr(), torch.cuda.current_stream().cuda_stream)
    assert rc == -1 and "at most 16515072 events" in L.gpu_phs_last_error_message().decode()
