"""The adjoint kernel against the oracle's forward-mode dual numbers and central finite
differences, and the torch.autograd wrapper."""
import numpy as np
import pytest

import oracle
from gpu_phs_amd import workloads as W

pytestmark = pytest.mark.gpu


def _setup(proc, n, seed, lo=0.02, E1=None, E2=None):
    import torch
    import gpu_phs_amd as G
    rng = np.random.default_rng(seed)
    counts = np.full(proc.n_channels, n // proc.n_channels)
    counts[: n - counts.sum()] += 1
    ch = np.repeat(np.arange(proc.n_channels), counts).astype(np.int32)
    x = lo + (1 - 2 * lo) * rng.random((n, proc.n_x))
    ph = G.ProcessHandle(proc, E1=E1, E2=E2)
    xt = torch.from_numpy(x).cuda()
    bufs, segs = [], []
    start = 0
    for c, k in enumerate(counts):
        b = G.EventBuffers.allocate(xt[start:start + k], proc.n_out)
        bufs.append(b)
        segs.append((ph, c, b))
        start += k
    plan = G.Plan(segs)
    plan.set_merge_below(0)   # every shape on its specialised adjoint, however few its events
    return rng, ch, x, ph, plan, bufs, counts


def _dual_grads(proc, x, ch, gp=None, gf=None, loss_mode=1, E1=None, E2=None):
    """d loss / d x from the oracle's dual numbers (K = n_x tangents)."""
    n = len(ch)
    K = proc.n_x
    xt = np.zeros((n, K, K))
    for k in range(K):
        xt[:, k, k] = 1
    d = oracle.forward_dual(proc, x, ch, xt, None, None, K, E1=E1, E2=E2)
    ref = oracle.forward(proc, x, ch, E1=E1, E2=E2)
    s = proc.sqrts ** 2
    if loss_mode == 1:
        S2 = (ref["p"][..., 1] ** 2 + ref["p"][..., 2] ** 2).sum(1)
        dS = 2 * (ref["p"][..., 1, None] * d["p_tan"][..., 1, :] + ref["p"][..., 2, None] * d["p_tan"][..., 2, :]).sum(1)
        g = d["factor_tan"] * (S2 / s)[:, None] + ref["factor"][:, None] * dS / s
    else:
        g = (gp[..., None] * d["p_tan"]).sum((1, 2)) + gf[:, None] * d["factor_tan"]
    g[~ref["ok"]] = 0
    return g, ref


def _dual_param_grads(proc, x, ch, ok, slots, gp=None, gf=None, loss_mode=1, E1=None, E2=None):
    """d loss / d (M, Gamma) of every (channel, bincode) slot from dual numbers."""
    s = proc.sqrts ** 2
    out = np.zeros((len(slots), 2))
    ref = oracle.forward(proc, x, ch, E1=E1, E2=E2)
    for i, (c, bc) in enumerate(slots):
        sel = (ch == c) & ok
        if not sel.any():
            continue
        for w in (0, 1):
            mt = np.zeros((proc.n_channels, proc.n_prt_out, 1))
            wt = np.zeros_like(mt)
            (mt if w == 0 else wt)[c, bc - 1, 0] = 1
            dd = oracle.forward_dual(proc, x[sel], ch[sel], None, mt, wt, 1,
                                     E1=None if E1 is None else E1[sel], E2=None if E2 is None else E2[sel])
            if loss_mode == 1:
                p = ref["p"][sel]
                S2 = (p[..., 1] ** 2 + p[..., 2] ** 2).sum(1)
                dS = 2 * (p[..., 1] * dd["p_tan"][..., 1, 0] + p[..., 2] * dd["p_tan"][..., 2, 0]).sum(1)
                out[i, w] = (dd["factor_tan"][:, 0] * S2 / s + ref["factor"][sel] * dS / s).sum()
            else:
                out[i, w] = ((gp[sel] * dd["p_tan"][..., 0]).sum((1, 2)) + gf[sel] * dd["factor_tan"][:, 0]).sum()
    return out


@pytest.mark.parametrize("name", ["c1", "c0", "c2", "2to6"])
@pytest.mark.parametrize("loss_mode", [1, 0])
def test_adjoint_matches_dual_numbers(name, loss_mode):
    import torch
    proc = {"c0": W.process_c0, "c1": W.process_c1, "c2": W.process_c2, "2to6": W.process_2to6}[name]()
    n = 3000 if name != "c2" else 6000
    rng, ch, x, ph, plan, bufs, counts = _setup(proc, n, 17)
    gp = rng.standard_normal((n, proc.n_out, 4)) if loss_mode == 0 else None
    gf = rng.standard_normal(n) if loss_mode == 0 else None
    gx = [torch.empty_like(b.x) for b in bufs]
    loss = [torch.empty(b.x.shape[0], dtype=torch.float64, device="cuda") for b in bufs]
    kw = {}
    if loss_mode == 0:
        gpt, gft = torch.from_numpy(gp).cuda(), torch.from_numpy(gf).cuda()
        kw = dict(g_p=list(torch.split(gpt, list(counts))), g_f=list(torch.split(gft, list(counts))))
    gparam = plan.adjoint(gx, loss=loss, loss_mode=loss_mode, **kw)
    torch.cuda.synchronize()
    # a second call gives the same bits: the reduction's order is fixed and its per-channel
    # completion counters are back at 0 after every launch
    first = gparam.clone()
    gparam = plan.adjoint(gx, loss=loss, loss_mode=loss_mode, **kw)
    torch.cuda.synchronize()
    assert torch.equal(first, gparam)
    g_gpu = torch.cat(gx).cpu().numpy()
    g_ref, ref = _dual_grads(proc, x, ch, gp, gf, loss_mode)
    ok = ref["ok"]
    scale = np.abs(g_ref).max(axis=1, keepdims=True) + 1e-300
    err = np.abs(g_gpu - g_ref) / scale
    print(name, loss_mode, "max rel err", err[ok].max())
    assert err[ok].max() < 1e-9
    assert (g_gpu[~ok] == 0).all()
    # forward outputs written by the adjoint kernel = the oracle's
    okg = torch.cat([b.ok for b in bufs]).cpu().numpy().astype(bool)
    assert np.array_equal(okg, ok)
    if loss_mode == 1:
        l_gpu = torch.cat(loss).cpu().numpy()
        S2 = (ref["p"][..., 1] ** 2 + ref["p"][..., 2] ** 2).sum(1)
        assert np.allclose(l_gpu[ok], (ref["factor"] * S2 / proc.sqrts ** 2)[ok], rtol=1e-12)
    # parameter gradients
    pr, pch, pbc = plan.param_map()
    slots = list(zip(pch.tolist(), pbc.tolist()))
    gpar_ref = _dual_param_grads(proc, x, ch, ok, slots, gp, gf, loss_mode)
    gpar = gparam.cpu().numpy().reshape(-1, 2)[: len(slots)]
    sc = np.abs(gpar_ref).max() + 1e-300
    print(name, loss_mode, "param max rel err", np.abs(gpar - gpar_ref).max() / sc)
    assert np.abs(gpar - gpar_ref).max() <= 1e-9 * sc


@pytest.mark.parametrize("name", ["c1", "c2", "2to6"])
def test_adjoint_per_event_beams_matches_dual_numbers(name):
    """Per-event beam energies: each event's sqrt(s), lab boost and map constants (and their
    derivatives with respect to M and Gamma) at its own beams, against the oracle's dual numbers
    with the same beams."""
    import torch
    proc = {"c1": W.process_c1, "c2": W.process_c2, "2to6": W.process_2to6}[name]()
    n = 3000 if name != "c2" else 6000
    rng, ch, x, ph, plan, bufs, counts = _setup(proc, n, 23)
    h2 = proc.sqrts / 2
    E1, E2 = h2 * rng.uniform(0.8, 1.3, n), h2 * rng.uniform(0.8, 1.3, n)
    e1t, e2t = torch.from_numpy(E1).cuda(), torch.from_numpy(E2).cuda()
    e1s, e2s = torch.split(e1t, list(counts)), torch.split(e2t, list(counts))
    for i in range(len(bufs)):
        plan.set_event_beams(i, e1s[i], e2s[i])
    gp = rng.standard_normal((n, proc.n_out, 4))
    gf = rng.standard_normal(n)
    gpt, gft = torch.from_numpy(gp).cuda(), torch.from_numpy(gf).cuda()
    kw = dict(g_p=list(torch.split(gpt, list(counts))), g_f=list(torch.split(gft, list(counts))))
    gx = [torch.empty_like(b.x) for b in bufs]
    print("per-event beams adjoint", name, flush=True)
    gparam = plan.adjoint(gx, loss_mode=0, **kw)
    torch.cuda.synchronize()
    g_gpu = torch.cat(gx).cpu().numpy()
    g_ref, ref = _dual_grads(proc, x, ch, gp, gf, 0, E1=E1, E2=E2)
    ok = ref["ok"]
    okg = torch.cat([b.ok for b in bufs]).cpu().numpy().astype(bool)
    assert np.array_equal(okg, ok) and ok.sum() > 200
    scale = np.abs(g_ref).max(axis=1, keepdims=True) + 1e-300
    err = np.abs(g_gpu - g_ref) / scale
    print(name, "max rel err", err[ok].max())
    assert err[ok].max() < 1e-9
    assert (g_gpu[~ok] == 0).all()
    # the forward outputs the adjoint kernel wrote: each event at its own beams
    p_gpu = torch.cat([b.p for b in bufs]).cpu().numpy()
    assert np.allclose(p_gpu[ok], ref["p"][ok], rtol=0, atol=1e-10 * proc.sqrts)
    pr, pch, pbc = plan.param_map()
    slots = list(zip(pch.tolist(), pbc.tolist()))
    gpar_ref = _dual_param_grads(proc, x, ch, ok, slots, gp, gf, 0, E1=E1, E2=E2)
    gpar = gparam.cpu().numpy().reshape(-1, 2)[: len(slots)]
    sc = np.abs(gpar_ref).max() + 1e-300
    print(name, "param max rel err", np.abs(gpar - gpar_ref).max() / sc)
    assert np.abs(gpar - gpar_ref).max() <= 1e-9 * sc


@pytest.mark.parametrize("case", ["friends4", "friends4-loss", "friends6-unequal", "friends4-event-beams",
                                  "friends6-event-beams"])
def test_adjoint_friends_matches_dual_numbers(case):
    """Branches whose polar axis is an incoming beam (friend signs -1 and +1, on the root and on
    inner branches): their boost depends on their parent boost through the beam's direction,
    so the adjoint of the parent boost has a second term.  Against the oracle's dual numbers
    through the same friend construction, with external cotangents and with the built-in loss,
    at equal, unequal and per-event beams; d loss / d x and d loss / d (M, Gamma)."""
    import torch
    n_out = 6 if "friends6" in case else 4
    proc = W.process_friends(n_out)
    n = 4000 if n_out == 4 else 12000
    lm = 1 if case.endswith("loss") else 0
    E1 = E2 = None
    if "unequal" in case:
        E1, E2 = 0.6 * proc.sqrts, proc.sqrts ** 2 / (4 * 0.6 * proc.sqrts)
    rng, ch, x, ph, plan, bufs, counts = _setup(proc, n, 41 + n_out, E1=E1, E2=E2)
    e1 = e2 = None
    if E1 is not None:
        e1, e2 = np.full(n, E1), np.full(n, E2)
    if "event-beams" in case:
        h2 = proc.sqrts / 2
        e1, e2 = h2 * rng.uniform(0.8, 1.3, n), h2 * rng.uniform(0.8, 1.3, n)
        e1s = torch.split(torch.from_numpy(e1).cuda(), list(counts))
        e2s = torch.split(torch.from_numpy(e2).cuda(), list(counts))
        for i in range(len(bufs)):
            plan.set_event_beams(i, e1s[i], e2s[i])
    gp = rng.standard_normal((n, proc.n_out, 4))
    gf = rng.standard_normal(n)
    kw = {}
    if lm == 0:
        gpt, gft = torch.from_numpy(gp).cuda(), torch.from_numpy(gf).cuda()
        kw = dict(g_p=list(torch.split(gpt, list(counts))), g_f=list(torch.split(gft, list(counts))))
    gx = [torch.empty_like(b.x) for b in bufs]
    print("friend adjoint", case, flush=True)
    gparam = plan.adjoint(gx, loss_mode=lm, **kw)
    torch.cuda.synchronize()
    g_gpu = torch.cat(gx).cpu().numpy()
    g_ref, ref = _dual_grads(proc, x, ch, gp, gf, lm, E1=e1, E2=e2)
    ok = ref["ok"]
    okg = torch.cat([b.ok for b in bufs]).cpu().numpy().astype(bool)
    assert np.array_equal(okg, ok) and ok.sum() > 200
    scale = np.abs(g_ref).max(axis=1, keepdims=True) + 1e-300
    err = np.abs(g_gpu - g_ref) / scale
    print(case, "max rel err", err[ok].max())
    assert err[ok].max() < 1e-9
    assert (g_gpu[~ok] == 0).all()
    p_gpu = torch.cat([b.p for b in bufs]).cpu().numpy()
    assert np.allclose(p_gpu[ok], ref["p"][ok], rtol=0, atol=1e-10 * proc.sqrts)
    pr, pch, pbc = plan.param_map()
    slots = list(zip(pch.tolist(), pbc.tolist()))
    gpar_ref = _dual_param_grads(proc, x, ch, ok, slots, gp, gf, lm, E1=e1, E2=e2)
    gpar = gparam.cpu().numpy().reshape(-1, 2)[: len(slots)]
    sc = np.abs(gpar_ref).max() + 1e-300
    print(case, "param max rel err", np.abs(gpar - gpar_ref).max() / sc)
    assert np.abs(gpar - gpar_ref).max() <= 1e-9 * sc


@pytest.mark.parametrize("n_out,n", [(8, 60000), (10, 600000)])
def test_adjoint_large_trees(n_out, n):
    """Trees of 8 and 10 final-state particles (save banks up to three levels deep, randoms and
    (M, Gamma) accumulators in LDS; the tape of a 2->10 tree takes 75 KB of a workgroup's LDS)
    against the oracle's dual numbers: d loss / d x over the first (up to 24) randoms, and
    d loss / d (M, Gamma)."""
    import torch
    proc = W.process_large(n_out)
    rng, ch, x, ph, plan, bufs, counts = _setup(proc, n, 31)
    gp = rng.standard_normal((n, proc.n_out, 4))
    gf = rng.standard_normal(n)
    gpt, gft = torch.from_numpy(gp).cuda(), torch.from_numpy(gf).cuda()
    kw = dict(g_p=list(torch.split(gpt, list(counts))), g_f=list(torch.split(gft, list(counts))))
    gx = [torch.empty_like(b.x) for b in bufs]
    print("adjoint 2 ->", n_out, flush=True)
    gparam = plan.adjoint(gx, loss_mode=0, **kw)
    torch.cuda.synchronize()
    g_gpu = torch.cat(gx).cpu().numpy()
    ref = oracle.forward(proc, x, ch)
    ok = ref["ok"]
    assert ok.sum() >= 100
    sel = np.nonzero(ok)[0]
    K = min(proc.n_x, 24)
    xt = np.zeros((sel.size, proc.n_x, K))
    for k in range(K):
        xt[:, k, k] = 1
    d = oracle.forward_dual(proc, x[sel], ch[sel], xt, None, None, K)
    g_ref = (gp[sel][..., None] * d["p_tan"]).sum((1, 2)) + gf[sel][:, None] * d["factor_tan"]
    scale = np.abs(g_ref).max(axis=1, keepdims=True) + 1e-300
    err = np.abs(g_gpu[sel][:, :K] - g_ref) / scale
    print(n_out, "max rel err", err.max())
    assert err.max() < 1e-9
    assert (g_gpu[~ok] == 0).all()
    pr, pch, pbc = plan.param_map()
    slots = list(zip(pch.tolist(), pbc.tolist()))
    gpar_ref = _dual_param_grads(proc, x, ch, ok, slots, gp, gf, 0)
    gpar = gparam.cpu().numpy().reshape(-1, 2)[: len(slots)]
    sc = np.abs(gpar_ref).max() + 1e-300
    print(n_out, "param max rel err", np.abs(gpar - gpar_ref).max() / sc)
    assert np.abs(gpar - gpar_ref).max() <= 1e-9 * sc


def test_adjoint_central_differences_c4():
    """The C4 gradient on the c1 trees against central differences (h = 1e-6) of the loss."""
    import torch
    proc = W.process_c1()
    n = 600
    rng, ch, x, ph, plan, bufs, counts = _setup(proc, n, 5, lo=0.05)
    gx = [torch.empty_like(b.x) for b in bufs]
    plan.adjoint(gx, loss_mode=1)
    g = torch.cat(gx).cpu().numpy()
    s = proc.sqrts ** 2

    def loss(xx):
        r = oracle.forward(proc, xx, ch)
        S2 = (r["p"][..., 1] ** 2 + r["p"][..., 2] ** 2).sum(1)
        return np.where(r["ok"], r["factor"] * S2 / s, 0.0), r["ok"]

    h = 1e-6
    _, ok0 = loss(x)
    fds, goods = [], []
    for k in range(proc.n_x):
        xp, xm = x.copy(), x.copy()
        xp[:, k] += h
        xm[:, k] -= h
        lp, okp = loss(xp)
        lm, okm = loss(xm)
        fds.append((lp - lm) / (2 * h))
        goods.append(ok0 & okp & okm)
    # one absolute scale for all randoms: the root's phi leaves sum pT^2 unchanged, so its
    # derivative is zero and its difference quotient is rounding noise (~eps * loss / h)
    sc = max(np.abs(fd[good]).max() for fd, good in zip(fds, goods))
    for k, (fd, good) in enumerate(zip(fds, goods)):
        assert np.allclose(g[good, k], fd[good], rtol=1e-4, atol=1e-6 * sc), k


def test_autograd_function_and_dot_product():
    import torch
    from gpu_phs_amd.autograd import PhaseSpace
    proc = W.process_c1()
    n = 2000
    counts = [700, 600, 700]
    rng = np.random.default_rng(9)
    x = torch.from_numpy(0.02 + 0.96 * rng.random((n, proc.n_x))).cuda().requires_grad_(True)
    M = torch.from_numpy(proc.map_masses.copy()).requires_grad_(True)
    Gm = torch.from_numpy(proc.map_widths.copy()).requires_grad_(True)
    ps = PhaseSpace(proc, counts)
    p, f, v, ok = ps(x, M, Gm)
    gp = torch.from_numpy(rng.standard_normal((n, proc.n_out, 4))).cuda()
    gf = torch.from_numpy(rng.standard_normal(n)).cuda()
    okf = ok.double()
    L = ((p * gp).sum((1, 2)) * okf).sum() + (f * gf * okf).sum()
    L.backward()
    ch = np.repeat(np.arange(3), counts).astype(np.int32)
    g_ref, ref = _dual_grads(proc, x.detach().cpu().numpy(), ch, gp.cpu().numpy(), gf.cpu().numpy(), 0)
    okn = ok.cpu().numpy()
    err = np.abs(x.grad.cpu().numpy() - g_ref) / (np.abs(g_ref).max(1, keepdims=True) + 1e-300)
    assert err[okn].max() < 1e-9
    # BW mass gradient of channel 0, branch 3 = (12): against dual numbers
    bw = [(0, 3), (0, 12), (1, 9), (1, 6), (2, 12)]
    ref_par = _dual_param_grads(proc, x.detach().cpu().numpy(), ch, okn, bw, gp.cpu().numpy(), gf.cpu().numpy(), 0)
    got = np.array([[M.grad[c, b - 1].item(), Gm.grad[c, b - 1].item()] for c, b in bw])
    assert np.allclose(got, ref_par, rtol=1e-9, atol=1e-12 * np.abs(ref_par).max())


@pytest.mark.parametrize("n_out", [2, 3, 4])
@pytest.mark.parametrize("loss_mode", [1, 0])
def test_shape_adjoint_identical_to_generic(n_out, loss_mode):
    """Every ordered tree of 2-4 final-state particles: the adjoint specialised for its shape
    gives bit for bit the generic adjoint's forward outputs and losses, and its gradients to
    1e-12 (the derivative code lets the compiler fuse products, which it does differently in
    the unrolled kernels)."""
    import torch
    proc = W.process_shapes(n_out)
    res = []
    for spec in (True, False):
        rng, ch, x, ph, plan, bufs, counts = _setup(proc, 12001, 5 + n_out)
        if not spec:
            plan.set_specialize(False)
        g_x = [torch.zeros_like(b.x) for b in bufs]
        loss = [torch.zeros(b.x.shape[0], dtype=torch.float64, device="cuda") for b in bufs]
        gp = [torch.from_numpy(rng.standard_normal((b.x.shape[0], proc.n_out, 4))).cuda() for b in bufs]
        gf = [torch.from_numpy(rng.standard_normal(b.x.shape[0])).cuda() for b in bufs]
        gparam = plan.adjoint(g_x, g_p=gp if loss_mode == 0 else None, g_f=gf if loss_mode == 0 else None,
                              loss=loss, loss_mode=loss_mode)
        torch.cuda.synchronize()
        cat = lambda ts: torch.cat(ts).cpu().numpy()
        res.append(dict(gx=cat(g_x), loss=cat(loss), gparam=gparam.cpu().numpy(), p=cat([b.p for b in bufs]),
                        f=cat([b.factor for b in bufs]), ok=cat([b.ok for b in bufs])))
    a, b = res
    assert a["ok"].sum() > 100
    for k in ("loss", "p", "f", "ok"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k
    for k in ("gx", "gparam"):
        scale = np.abs(b[k]).max(axis=-1, keepdims=True) if k == "gx" else np.abs(b[k]).max()
        assert np.all(np.abs(a[k] - b[k]) <= 1e-12 * scale), (k, np.abs(a[k] - b[k]).max())


@pytest.mark.parametrize("n_out", [2, 3, 4])
def test_shape_adjoint_identical_to_generic_event_beams(n_out):
    """Per-event beam energies (loss mode 0): the multi-shape adjoint of the trees of 2-4
    final-state particles gives bit for bit the generic adjoint's forward outputs, and its
    gradients to 1e-12 (it takes the derivatives of the maps' constants at the event's sqrt(s)
    after the sweep, a different order of the same sums)."""
    import torch
    proc = W.process_shapes(n_out)
    res = []
    for spec in (True, False):
        rng, ch, x, ph, plan, bufs, counts = _setup(proc, 12001, 15 + n_out)
        if not spec:
            plan.set_specialize(False)
        h2 = proc.sqrts / 2
        for i, b in enumerate(bufs):
            k = b.x.shape[0]
            plan.set_event_beams(i, torch.from_numpy(h2 * rng.uniform(0.8, 1.3, k)).cuda(),
                                 torch.from_numpy(h2 * rng.uniform(0.8, 1.3, k)).cuda())
        g_x = [torch.zeros_like(b.x) for b in bufs]
        gp = [torch.from_numpy(rng.standard_normal((b.x.shape[0], proc.n_out, 4))).cuda() for b in bufs]
        gf = [torch.from_numpy(rng.standard_normal(b.x.shape[0])).cuda() for b in bufs]
        print("event-beams adjoint, n_out", n_out, "specialised" if spec else "generic", flush=True)
        gparam = plan.adjoint(g_x, g_p=gp, g_f=gf, loss_mode=0)
        torch.cuda.synchronize()
        cat = lambda ts: torch.cat(ts).cpu().numpy()
        res.append(dict(gx=cat(g_x), gparam=gparam.cpu().numpy(), p=cat([b.p for b in bufs]),
                        f=cat([b.factor for b in bufs]), v=cat([b.volume for b in bufs]),
                        ok=cat([b.ok for b in bufs])))
    a, b = res
    assert a["ok"].sum() > 100
    for k in ("p", "f", "v", "ok"):
        assert np.array_equal(a[k], b[k], equal_nan=True), k
    for k in ("gx", "gparam"):
        scale = np.abs(b[k]).max(axis=-1, keepdims=True) if k == "gx" else np.abs(b[k]).max()
        assert np.all(np.abs(a[k] - b[k]) <= 1e-12 * scale), (k, np.abs(a[k] - b[k]).max())


def test_adjoint_kernels_do_not_depend_on_merge_below():
    """The adjoint forms its own runs by tree shape: with the forward pass's default merge_below
    (which folds these small segments into its generic launch) the adjoint still runs the
    shape-specialised code, and gives the same bits as with merge_below 0."""
    import torch
    proc = W.process_shapes(4)
    res = []
    for mb in (None, 0):
        rng, ch, x, ph, plan, bufs, counts = _setup(proc, 6001, 31)
        if mb is None:
            plan.set_merge_below(1 << 18)
        g_x = [torch.zeros_like(b.x) for b in bufs]
        loss = [torch.zeros(b.x.shape[0], dtype=torch.float64, device="cuda") for b in bufs]
        print("adjoint, merge_below", mb, flush=True)
        gparam = plan.adjoint(g_x, loss=loss, loss_mode=1)
        torch.cuda.synchronize()
        res.append((torch.cat(g_x).cpu().numpy(), torch.cat(loss).cpu().numpy(), gparam.cpu().numpy()))
    for a, b in zip(*res):
        assert np.array_equal(a, b)


def test_adjoint_channel_without_events_gets_zero_gradients():
    """A segment of no events makes a plan channel without tiles: its (M, Gamma) slots are set to
    0 (every other slot is the sum over its channel's units, as in a plan without that segment),
    and two calls in a row, the second on fewer events, give the gradients of their own inputs
    (the reduction's counters are back at 0 after each call)."""
    import torch
    import gpu_phs_amd as G
    proc = W.process_c1()
    ph = G.ProcessHandle(proc)
    rng = np.random.default_rng(5)
    xs = [torch.from_numpy(0.02 + 0.96 * rng.random((n, proc.n_x))).cuda() for n in (3000, 0, 5000)]
    bufs = [G.EventBuffers.allocate(x, proc.n_out) for x in xs]
    plan = G.Plan([(ph, c, b) for c, b in zip((0, 1, 2), bufs)])
    gx = [torch.empty_like(b.x) for b in bufs]
    gparam = torch.full((plan.n_params(),), float("nan"), dtype=torch.float64, device="cuda")
    print("adjoint with an empty channel", flush=True)
    plan.adjoint(gx, loss_mode=1, gparam=gparam)
    torch.cuda.synchronize()
    pr, pch, pbc = plan.param_map()
    g = gparam.cpu().numpy().reshape(-1, 2)
    assert (g[pch == 1] == 0).all()
    assert np.isfinite(g).all()
    # the same channels without the empty segment
    ref_plan = G.Plan([(ph, 0, bufs[0]), (ph, 2, bufs[2])])
    gx2 = [torch.empty_like(bufs[0].x), torch.empty_like(bufs[2].x)]
    gref = ref_plan.adjoint(gx2, loss_mode=1).cpu().numpy().reshape(-1, 2)
    _, rch, rbc = ref_plan.param_map()
    for c in (0, 2):
        mine = {int(b): g[i] for i, (cc, b) in enumerate(zip(pch, pbc)) if cc == c}
        theirs = {int(b): gref[i] for i, (cc, b) in enumerate(zip(rch, rbc)) if cc == c}
        assert mine.keys() == theirs.keys()
        for b in mine:
            assert np.array_equal(mine[b], theirs[b]), (c, b)
    # a second call on fewer events of the same plan layout: its own gradients
    small = [torch.from_numpy(0.02 + 0.96 * rng.random((n, proc.n_x))).cuda() for n in (700, 0, 900)]
    sb = [G.EventBuffers.allocate(x, proc.n_out) for x in small]
    p2 = G.Plan([(ph, c, b) for c, b in zip((0, 1, 2), sb)])
    sgx = [torch.empty_like(b.x) for b in sb]
    first = p2.adjoint(sgx, loss_mode=1).clone()
    again = p2.adjoint(sgx, loss_mode=1)
    torch.cuda.synchronize()
    assert torch.equal(first, again)


def test_adjoint_reduction_over_many_units_per_channel():
    """Channels of more than 8,192 64-event units reduce through 16 range sums per channel
    (fewer units: one workgroup per channel): the (M, Gamma) gradients of 537,600 events per
    channel equal the sum of those of its four quarters, each reduced in one range, to 1e-12."""
    import torch
    import gpu_phs_amd as G
    proc = W.process_c1()
    ph = G.ProcessHandle(proc)
    rng = np.random.default_rng(8)
    per = 8400 * 64
    xs = [torch.from_numpy(0.02 + 0.96 * rng.random((per, proc.n_x))).cuda() for _ in range(proc.n_channels)]
    bufs = [G.EventBuffers.allocate(x, proc.n_out) for x in xs]
    plan = G.Plan([(ph, c, b) for c, b in enumerate(bufs)])
    print("adjoint over 8,400 units per channel", flush=True)
    full = plan.adjoint([torch.empty_like(b.x) for b in bufs], loss_mode=1).cpu().numpy()
    q = per // 4
    parts = 0
    for k in range(4):
        qb = [G.EventBuffers.allocate(x[k * q:(k + 1) * q], proc.n_out) for x in xs]
        qp = G.Plan([(ph, c, b) for c, b in enumerate(qb)])
        assert list(zip(*qp.param_map())) == list(zip(*plan.param_map()))
        parts = parts + qp.adjoint([torch.empty_like(b.x) for b in qb], loss_mode=1).cpu().numpy()
    torch.cuda.synchronize()
    sc = np.abs(full).max()
    print("max rel diff", np.abs(full - parts).max() / sc)
    assert np.abs(full - parts).max() <= 1e-12 * sc
