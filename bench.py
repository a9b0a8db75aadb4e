"""Benchmark: events/s of the phase-space generator on one MI355X (FP64).

Default workload: the batched uneven mix (every channel of the 2->2, 2->4 double-resonance and
120-channel 2->4 processes plus sixteen 2->6 trees, Dirichlet(0.5) event counts, 5e7 events),
the largest configuration, run as the plan's batched launches.  One step = one forward pass over
all 5e7 events, randoms resident in HBM, outputs (momenta, factor, volume, ok) written to HBM.
The JSON line also carries the plain per-segment launches, the other workloads, the adjoint,
the inverse mapping and the Whizard entry point as extra fields, the HBM and FP64 rooflines of
the forward kernels and the reference's own CPU algorithm timed on the host's cores.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3] [--no-extra]
                    [--tuned '<json of Plan.tuning()>' | --config block,ept,grid,order,wpe,lds | --autotune]

The c3 mix runs the launch tuning C3_TUNING (the autotuner's choice, pinned so that the line and
the committed counter profiles describe the same kernels); the autotuner's own result on the same
mix is reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector, dense


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(wl, torch, G):
    """Device randoms (seeded U(0,1) per segment) and output buffers of a workload."""
    handles = [G.ProcessHandle(p) for p in wl.processes]
    gen = torch.Generator(device="cuda")
    gen.manual_seed(wl.seed & 0x7FFFFFFFFFFFFFFF)
    segs = []
    for s in wl.segments:
        ph = handles[s.process]
        x = torch.rand((s.n_events, ph.n_x), dtype=torch.float64, device="cuda", generator=gen)
        segs.append((ph, s.channel, G.EventBuffers.allocate(x, ph.n_out)))
    return handles, segs


def time_steps(fn, steps, warmup, torch):
    """Median-free wall timing of `steps` calls bracketed by synchronize, plus HIP-event time
    of the same region on the launch stream (torch's current stream)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for _ in range(steps):
        fn()
    end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, start.elapsed_time(end) / 1e3


def run_workload(name, steps, warmup, torch, G, W, tune=True, fixed=None, tuning=None, specialize=True):
    """fixed: one launch shape for both tile classes; tuning: a Plan.tuning() dict to apply;
    neither: the autotuner chooses.  specialize=False: the generic kernel for every tree."""
    wl = W.workload(name)
    handles, segs = make_inputs(wl, torch, G)
    plan = G.Plan(segs)
    if not specialize:
        plan.set_specialize(False)
    if fixed:
        plan.set_config(**fixed)
        tuned = ""
    elif tuning:
        plan.apply_tuning(tuning)
        tuned = "applied: " + json.dumps(tuning)
    else:
        tuned = plan.autotune(reps=3) if tune else ""
    torch.cuda.synchronize()
    # the batched launches submitted from a hipGraph and directly; a graph replay costs a few
    # microseconds more than a direct launch, which shows on the smallest workloads
    wall, ev_s = time_steps(plan.run_graph, steps, warmup, torch)
    launch = "one batched launch per tile class, replayed from a hipGraph"
    d_wall, d_ev = time_steps(plan.run, steps, warmup, torch)
    if d_ev < ev_s:
        wall, ev_s = d_wall, d_ev
        launch = "one batched launch per tile class, submitted directly"
    plain_wall, plain_ev = time_steps(plan.run_plain, max(1, steps // 2), warmup, torch)
    n = wl.n_events
    res = dict(workload=name, n_events=n, n_segments=len(segs),
               batched_ev_s=n * steps / wall, batched_kernel_ms=ev_s / steps * 1e3,
               plain_ev_s=n * max(1, steps // 2) / plain_wall, plain_ms=plain_ev / max(1, steps // 2) * 1e3,
               algorithmic_bytes=wl.algorithmic_bytes(), config=plan.config(), tuning=plan.tuning(),
               tune=tuned[:200], wall_s=wall, event_s=ev_s, launch=launch)
    res["batched_vs_plain"] = res["batched_ev_s"] / res["plain_ev_s"]
    return res, plan, segs, wl


def run_adjoint(steps, warmup, torch, G, W, specialize=True):
    """c4: forward + reverse mode on the c1 trees, 1e6 events, loss = sum ok * factor * sum pT^2 / s.
    One adjoint launch runs the forward pass and writes the per-event loss and d loss / d (x, M,
    Gamma): 64 B read and 72 B written per event, the c4 configuration (randoms in, their
    gradient and the loss out).  The same launch that also writes the forward outputs (momenta,
    factor, volume, ok: 145 B more per event) is timed beside it, and so is loss mode 0 (the
    cotangents of momenta and factor given), with the fixed beams and with per-event beams."""
    wl = W.workload("c4")
    handles, segs = make_inputs(wl, torch, G)
    plan = G.Plan(segs)
    if not specialize:
        plan.set_specialize(False)
    plan.autotune(reps=3)
    gx = [torch.empty_like(b.x) for _, _, b in segs]
    loss = [torch.empty(b.x.shape[0], dtype=torch.float64, device="cuda") for _, _, b in segs]
    gparam = plan.adjoint(gx, loss=loss, loss_mode=1)
    fw_wall, fw_ev = time_steps(plan.run_graph, steps, warmup, torch)
    ad_wall, ad_ev = time_steps(lambda: plan.adjoint(gx, loss=loss, loss_mode=1, gparam=gparam, write_forward=False),
                                steps, warmup, torch)
    am_wall, am_ev = time_steps(lambda: plan.adjoint(gx, loss=loss, loss_mode=1, gparam=gparam, write_forward=True),
                                steps, warmup, torch)
    # loss mode 0: seeded cotangents of the momenta and factors in (another 136 B read per
    # event), first with the fixed beams, then with beam energies that differ per event
    gen = torch.Generator(device="cuda")
    gen.manual_seed(11)
    gp = [torch.rand(b.p.shape, dtype=torch.float64, device="cuda", generator=gen) for _, _, b in segs]
    gf = [torch.rand(b.factor.shape, dtype=torch.float64, device="cuda", generator=gen) for _, _, b in segs]
    mode0 = lambda: plan.adjoint(gx, g_p=gp, g_f=gf, loss_mode=0, gparam=gparam, write_forward=False)
    m0_wall, m0_ev = time_steps(mode0, steps, warmup, torch)
    for i, (_, _, b) in enumerate(segs):
        k = b.x.shape[0]
        plan.set_event_beams(i, 250.0 * (0.8 + 0.5 * torch.rand(k, dtype=torch.float64, device="cuda", generator=gen)),
                             250.0 * (0.8 + 0.5 * torch.rand(k, dtype=torch.float64, device="cuda", generator=gen)))
    fb_wall, fb_ev = time_steps(plan.run_graph, steps, warmup, torch)
    pb_wall, pb_ev = time_steps(mode0, steps, warmup, torch)
    n = wl.n_events
    return dict(n_events=n, forward_ev_s=n * steps / fw_wall, fwd_adjoint_ev_s=n * steps / ad_wall,
                forward_ms=fw_ev / steps * 1e3, fwd_adjoint_ms=ad_ev / steps * 1e3,
                fwd_adjoint_over_forward=ad_wall / fw_wall,
                fwd_adjoint_with_outputs_ms=am_ev / steps * 1e3,
                fwd_adjoint_with_outputs_over_forward=am_wall / fw_wall,
                loss_mode0_ms=m0_ev / steps * 1e3, loss_mode0_over_forward=m0_wall / fw_wall,
                event_beams_forward_ms=fb_ev / steps * 1e3,
                event_beams_loss_mode0_ms=pb_ev / steps * 1e3,
                event_beams_loss_mode0_over_forward=pb_wall / fb_wall,
                loss=float(sum(l.sum().item() for l in loss)))


def run_inverse(name, steps, warmup, torch, G, W, scale=1.0, block=0):
    """All-channel inverse mapping of a workload's events (after one forward pass): every
    event's randoms and factor in every channel of its process, one launch."""
    wl = W.workload(name, scale=scale)
    handles, segs = make_inputs(wl, torch, G)
    plan = G.Plan(segs)
    plan.run()
    proc = wl.processes[0]
    n = wl.n_events
    f = torch.empty((n, proc.n_channels), dtype=torch.float64, device="cuda")
    xo = torch.empty((proc.n_channels, n, proc.n_x), dtype=torch.float64, device="cuda")
    plan.inverse(factors=f, x_out=xo, block=block)
    wall, ev_s = time_steps(lambda: plan.inverse(factors=f, x_out=xo, block=block), steps, warmup, torch)
    nc = proc.n_channels
    # reads: momenta, the own randoms and factor; writes: every channel's factor and randoms
    bytes_ = n * (32 * proc.n_out + 8 * proc.n_x + 8) + n * nc * (8 + 8 * proc.n_x)
    ms = ev_s / steps * 1e3
    return dict(n_events=n, channels=nc, kernel_ms=ms, events_s=n * steps / wall,
                pairs_s=n * nc * steps / wall, bytes_per_launch=bytes_,
                hbm_frac=bytes_ / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                block=block or ("library default: 256 for up to 16 channels, else 128"))


def run_event_beams(steps, torch, G, W):
    """The c1 workload with beam energies that differ per event (each event its own sqrt(s),
    lab boost and sqrt(s)-dependent map constants)."""
    wl = W.workload("c1")
    handles, segs = make_inputs(wl, torch, G)
    plan = G.Plan(segs)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(7)
    beams = []
    for i, (_, _, b) in enumerate(segs):
        n = b.x.shape[0]
        e1 = 250.0 * (0.8 + 0.5 * torch.rand(n, dtype=torch.float64, device="cuda", generator=gen))
        e2 = 250.0 * (0.8 + 0.5 * torch.rand(n, dtype=torch.float64, device="cuda", generator=gen))
        beams.append((e1, e2))
        plan.set_event_beams(i, e1, e2)
    tuned = plan.autotune(reps=3)
    # replayed from a hipGraph and submitted directly, the faster of the two (as for the
    # fixed-beam workloads)
    wall, ev = time_steps(plan.run_graph, steps, 5, torch)
    launch = "replayed from a hipGraph"
    d_wall, d_ev = time_steps(plan.run, steps, 5, torch)
    if d_ev < ev:
        wall, ev, launch = d_wall, d_ev, "submitted directly"
    n = wl.n_events
    ms = ev / steps * 1e3
    bytes_ = wl.algorithmic_bytes() + 16 * n   # the beam energies are read as well
    return dict(n_events=n, kernel_ms=ms, events_s=n * steps / wall, hbm_frac=bytes_ / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                launch=launch, tune_summary=tuned.split(" c0:")[0])


def run_grid(steps, torch, G, W):
    """Randoms of the c1 workload (1e7 events x 8) drawn on the device from adaptive grids
    (Philox4x32-10 + inverse distribution function), and one grid update (histogram of the
    accepted events)."""
    wl = W.workload("c1")
    handles, segs = make_inputs(wl, torch, G)
    plan = G.Plan(segs)
    grid = G.Grid(handles[0])
    w = np.random.default_rng(5).random(grid.shape) + 0.05
    grid.set_weights(w)
    wall, ev = time_steps(lambda: grid.generate(plan, 42), steps, 2, torch)
    plan.run()
    t0 = time.perf_counter()
    grid.update(plan)
    upd = time.perf_counter() - t0
    n = wl.n_events
    bytes_ = n * handles[0].n_x * 8
    ms = ev / steps * 1e3
    return dict(n_events=n, generate_ms=ms, randoms_s=n * handles[0].n_x / (ms / 1e3),
                hbm_frac=bytes_ / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, update_ms=upd * 1e3)


def cpu_inverse_baseline(name, scale, W):
    """The oracle's inverse mapping (a C++ restatement of the reference's, single thread) on a
    bounded sample of the workload's events: (event, channel) pairs per second."""
    import oracle
    wl = W.workload(name, scale=scale)
    proc = wl.processes[0]
    ch = np.concatenate([np.full(s.n_events, s.channel, np.int32) for s in wl.segments])
    x = np.random.default_rng(wl.seed).random((ch.shape[0], proc.n_x))
    p = oracle.forward(proc, x, ch)["p"]
    t0 = time.perf_counter()
    reps = 0
    while time.perf_counter() - t0 < 3.0 or reps == 0:
        oracle.inverse(proc, p)
        reps += 1
    dt = time.perf_counter() - t0
    pairs = ch.shape[0] * proc.n_channels * reps
    return dict(value=pairs / dt, unit="(event, channel) pairs/s", cores=1, kind="port",
                sample=f"{ch.shape[0]} events of {name} x {proc.n_channels} channels x {reps} passes, {dt:.1f} s")


def run_whizard_abi(n, reps, G, W):
    """The Whizard entry point end to end (host arrays in and out, the reference's call
    sequence), c1 trees: forward pass, all-channel inverse mapping, PCIe copies both ways."""
    import ctypes
    import numpy as np
    proc = W.process_c1()
    L = G.lib()
    ci = lambda v: ctypes.byref(ctypes.c_int(v))
    cp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    nc = proc.n_channels
    L.c_whizard_set_particle_structure(ci(nc), ci(nc), ci(1), ci(proc.n_x), ci(proc.n_in), ci(proc.n_out))
    flv = np.ascontiguousarray(proc.flv_masses)
    L.c_whizard_set_flavors(cp(flv), cp(np.zeros_like(flv)))
    L.c_whizard_init_mappings()
    L.c_whizard_init_tree_structures()
    keep = []
    for c in range(nc):
        a = [np.ascontiguousarray(t[c]) for t in (proc.map_ids, proc.map_masses, proc.map_widths, proc.d1, proc.d2,
                                                  proc.has_children)] + [np.zeros(proc.n_prt, np.int32)]
        keep.append(a)
        L.c_whizard_fill_mapping(ci(c + 1), cp(a[0]), cp(a[1]), cp(a[2]))
        L.c_whizard_fill_tree_structure(ci(c + 1), cp(a[3]), cp(a[4]), cp(a[5]), cp(a[6]))
    L.c_whizard_init_gpu_phs()
    counts = [n // 3, n // 3, n - 2 * (n // 3)]
    limits = (ctypes.c_int * 3)(*np.cumsum(counts).tolist())
    L.c_whizard_init_channel_ids(ci(n), ci(3), limits)
    x = np.random.default_rng(1).random((n, proc.n_x))
    e1, e2 = np.full(n, proc.sqrts / 2), np.full(n, proc.sqrts / 2)
    factors, volumes = np.zeros(n * nc), np.zeros(n)
    oks, p, x_out = np.zeros(n, np.bool_), np.zeros((n, proc.n_out, 4)), np.zeros(n * nc * proc.n_x)
    flag = ctypes.c_bool(True)
    call = lambda: L.c_whizard_gen_phs_from_x_gpu(cp(e1), cp(e2), ctypes.byref(flag), ci(n), ci(nc), ci(proc.n_x),
                                                  cp(x), cp(factors), cp(volumes), cp(oks), cp(p), cp(x_out))
    call()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    dt = (time.perf_counter() - t0) / reps
    if L.gpu_phs_last_error() != 0:
        raise RuntimeError(L.gpu_phs_last_error_message().decode())
    L.c_whizard_finalize_gpu_phs()
    host_bytes = x.nbytes + e1.nbytes + e2.nbytes + factors.nbytes + volumes.nbytes + oks.nbytes + p.nbytes + x_out.nbytes
    return dict(n_events=n, channels=nc, ms_per_call=dt * 1e3, events_s=n / dt, host_bytes_per_call=host_bytes,
                host_gb_s=host_bytes / dt / 1e9)


def cpu_threads():
    """Host threads for the CPU baseline: the cores this process may run on
    (sched_getaffinity), limited by the cgroup CPU quota where one is set and by 16, the CPU share
    of one GPU on a shared GPU host.  Returns (threads, affinity cores, quota cores or None)."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
        if q != "max":
            quota = max(1, -(-int(q) // int(period)))
    except (OSError, ValueError):
        pass
    return max(1, min(aff, quota or aff, 16)), aff, quota


def _process_samples(wl):
    """Per process of a workload: (process, channels, randoms) of its segments, randoms seeded."""
    rng = np.random.default_rng(wl.seed)
    out = []
    for pi, proc in enumerate(wl.processes):
        ch = np.concatenate([np.full(s.n_events, s.channel, np.int32) for s in wl.segments if s.process == pi] or
                            [np.zeros(0, np.int32)])
        if ch.size:
            out.append((proc, ch, rng.random((ch.shape[0], proc.n_x))))
    return out


def cpu_baseline(wl_name, W, threads, scale, budget_s):
    """The reference's own phs_cpu.cpp (oracle/_ref build; the oracle's restatement where that is
    absent) on a bounded sample of the workload -- the same processes, channels and event shares
    at `scale` of its size -- split over `threads` host threads (contiguous event ranges), repeated
    until about `budget_s` seconds of generation time."""
    import oracle
    wl = W.workload(wl_name, scale=scale)
    samples = _process_samples(wl)
    kind = "reference" if oracle.ref_available() else "port"
    total_t, total_n, reps = 0.0, 0, 0
    while (total_t < budget_s and reps < 40) or reps == 0:
        for proc, ch, x in samples:
            if kind == "reference":
                sec, _ = oracle.ref_time_threads(proc, x, ch, threads) if threads > 1 else oracle.ref_time(proc, x, ch)
            else:
                t0 = time.perf_counter()
                oracle.forward(proc, x, ch)
                sec = time.perf_counter() - t0
            total_t += sec
            total_n += ch.shape[0]
        reps += 1
    n = sum(ch.shape[0] for _, ch, _ in samples)
    cores, aff, quota = cpu_threads()
    return dict(value=total_n / total_t, unit="events/s", cores=threads if kind == "reference" else 1, kind=kind,
                sample=f"{n} events of {wl_name} ({scale:g} of it: same processes, channels and shares, seeded U(0,1)) "
                       f"x {reps} passes, {threads} thread(s) over contiguous event ranges, {total_t:.1f} s; "
                       f"host: {aff} cores in the affinity mask, cgroup quota {quota or 'none'}; "
                       f"reference phs_cpu.cpp built -O3")


# The launch tuning of the c3 mix that the headline line runs and the committed profiles were
# taken with, confirmed by interleaved repeats of the shapes around it (class 0: 128-thread
# workgroups, one event per thread, so one workgroup per tile of one sub-tile and the one-pass
# kernels at the compiler's register budget, four waves per SIMD; class 1: 64 threads, one event
# per thread, its one-pass kernels at three waves per SIMD; the classes share the GPU).  --autotune tunes
# instead, and the full run reports the autotuner's own result beside it (extra.c3_autotuned).
C3_TUNING = {"classes": [{"block": 128, "ept": 1, "grid": 0, "waves_per_eu": 0, "lds": 0},
                         {"block": 64, "ept": 1, "grid": 0, "waves_per_eu": 0, "lds": 0}],
             "order": 0, "serial": 0, "merge_below": 262144, "streams": 1}


def measured_counters(workload):
    """The committed counter profile of the forward kernels of one step of `workload`
    (profiles/forward_counters.json): HBM bytes per step from FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch, summed over the step's dispatches, each scaled by what the counter reads for a plain
    device copy), FP64 operations per step from the FP64 VALU instruction counters (64 lanes x
    (2 FMA + MUL + ADD + transcendental)), and the launch tuning they were taken with; None if
    the profile is of another workload."""
    path = os.path.join(ROOT, "profiles", "forward_counters.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("workload") != workload:
        return None
    c = t["per_step"]
    traffic = (c["FETCH_SIZE"] * t["fetch_scale"] + c["WRITE_SIZE"] * t["write_scale"]) * 1024
    flops = 64 * (2 * c["SQ_INSTS_VALU_FMA_F64"] + c["SQ_INSTS_VALU_MUL_F64"] + c["SQ_INSTS_VALU_ADD_F64"] +
                  c["SQ_INSTS_VALU_TRANS_F64"])
    return dict(traffic=traffic, flops=flops, tuning=t.get("tuning"), kernel_ms=t.get("kernel_ms_per_step"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--tuned", default="",
                    help="apply this launch tuning (JSON of Plan.tuning(), as the bench line reports it in "
                         "config.tuned); c3 runs C3_TUNING unless this, --config or --autotune is given")
    ap.add_argument("--autotune", action="store_true",
                    help="let the autotuner choose the tuning of the workload (the default for workloads "
                         "other than c3)")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-specialize", action="store_true",
                    help="run every tree on the generic forward kernel (no shape-specialised kernels)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", default="",
                    help="fixed launch config block,ept,grid,order,waves_per_eu,lds instead of autotuning")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only run the forward + adjoint workload (c4) and print its figures")
    ap.add_argument("--inverse-only", action="store_true",
                    help="only run the all-channel inverse mapping of c1 and c2 and print its figures")
    ap.add_argument("--grid-only", action="store_true",
                    help="only run the adaptive-grid generation and update on c1 and print the figures")
    ap.add_argument("--whizard-only", action="store_true",
                    help="only run the Whizard entry point end to end on 1e6 c1 events and print the figures")
    ap.add_argument("--beams-only", action="store_true",
                    help="only run c1 with per-event beam energies and print the figures")
    ap.add_argument("--forward-only", action="store_true",
                    help="only run the batched forward step of the workload (direct launches, --warmup + "
                         "--steps of them, with --tuned or the autotuner's tuning) and print its figures; "
                         "for counter profiles, whose per-step figures are the totals over all its dispatches "
                         "divided by warmup + steps")
    ap.add_argument("--sweep-schedule", action="store_true",
                    help="with --forward-only: time the step for merge_below in {0, 2^14, 2^16, 2^18, 2^20} x "
                         "classes shared / back to back, 3 interleaved repeats of --steps steps each, on top of "
                         "--tuned (or the autotuner's tuning), and print the times")
    ap.add_argument("--launch-modes", action="store_true",
                    help="with --forward-only: time the step submitted directly, replayed from a hipGraph, with "
                         "persistent grids of 2, 4 and 8 workgroups per CU (per 256 threads) in both classes, and "
                         "as plain per-segment launches; 3 interleaved repeats of --steps steps each")
    ap.add_argument("--calibrate-copy", type=float, default=0.0,
                    help="only run a plain device-to-device copy of this many GiB (3 times with torch's copy, "
                         "twice with the library's streaming-store copy), to scale the FETCH_SIZE / WRITE_SIZE "
                         "counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs on a single MI355X (got --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import gpu_phs_amd as G
    from gpu_phs_amd import workloads as W
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(2)
    G.lib()
    if args.calibrate_copy > 0:
        n = int(args.calibrate_copy * 2 ** 30) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        torch.cuda.empty_cache()
        # the library's copy kernel (stream_copy), whose stores are streaming stores like the
        # generator's: two dispatches of the same size
        G.copy_bandwidth(n * 8, 1)
        print(json.dumps({"calibrate_copy_bytes": n * 8, "fetch_expected_kib": n * 8 / 1024,
                          "write_expected_kib": n * 8 / 1024}))
        return
    if args.adjoint_only:
        print(json.dumps(run_adjoint(max(5, args.steps), args.warmup, torch, G, W, specialize=not args.no_specialize)))
        return
    if args.grid_only:
        print(json.dumps(run_grid(max(3, args.steps), torch, G, W)))
        return
    if args.whizard_only:
        print(json.dumps(run_whizard_abi(1_000_000, max(3, args.steps // 4), G, W)))
        return
    if args.beams_only:
        print(json.dumps(run_event_beams(max(3, args.steps), torch, G, W)))
        return
    if args.inverse_only:
        r = {"c1": run_inverse("c1", max(3, args.steps // 2), 1, torch, G, W),
             "c2": run_inverse("c2", max(3, args.steps // 2), 1, torch, G, W, scale=0.1)}
        print(json.dumps(r))
        return
    if args.forward_only:
        wl = W.workload(args.workload)
        handles, segs = make_inputs(wl, torch, G)
        plan = G.Plan(segs)
        if args.tuned:
            plan.apply_tuning(json.loads(args.tuned))
        elif args.workload == "c3" and not args.autotune:
            plan.apply_tuning(C3_TUNING)
        else:
            plan.autotune(reps=3)
        torch.cuda.synchronize()
        if args.launch_modes:
            base, times = plan.tuning(), {}
            n_cu = torch.cuda.get_device_properties(0).multi_processor_count
            modes = [("direct", None, "run"), ("graph", None, "run_graph"), ("plain", None, "run_plain")]
            for k in (2, 4, 8):
                modes.append((f"persistent {k}/CU direct", k, "run"))
                modes.append((f"persistent {k}/CU graph", k, "run_graph"))
            for _ in range(3):
                for name, k, fn in modes:
                    t = json.loads(json.dumps(base))
                    if k:
                        for lc in t["classes"]:
                            lc["grid"] = n_cu * k * (256 // lc["block"])
                    plan.apply_tuning(t)
                    torch.cuda.synchronize()
                    _, ev = time_steps(getattr(plan, fn), args.steps, args.warmup, torch)
                    times.setdefault(name, []).append(ev / args.steps * 1e3)
            print(json.dumps({"workload": args.workload, "base_tuning": base, "n_cu": n_cu,
                              "ms_per_step": {k: sorted(v) for k, v in times.items()}}))
            return
        if args.sweep_schedule:
            base, times = plan.tuning(), {}
            for _ in range(3):
                for mb in (0, 1 << 14, 1 << 16, 1 << 18, 1 << 20):
                    for serial in (0, 1):
                        plan.apply_tuning(dict(base, merge_below=mb, serial=serial))
                        torch.cuda.synchronize()
                        _, ev = time_steps(plan.run, args.steps, args.warmup, torch)
                        times.setdefault(f"merge_below={mb} serial={serial}", []).append(ev / args.steps * 1e3)
            print(json.dumps({"workload": args.workload, "base_tuning": base,
                              "ms_per_step": {k: sorted(v) for k, v in times.items()}}))
            return
        wall, ev_s = time_steps(plan.run, args.steps, args.warmup, torch)
        print(json.dumps({"workload": args.workload, "n_events": wl.n_events, "steps": args.steps,
                          "warmup": args.warmup, "tuning": plan.tuning(),
                          "kernel_ms_per_step": ev_s / args.steps * 1e3,
                          "events_s": wl.n_events * args.steps / wall}))
        return
    fixed = None
    if args.config:
        v = [int(t) for t in args.config.split(",")]
        fixed = dict(zip(("block", "ept", "grid", "order", "waves_per_eu", "lds"), v))
    tuning = json.loads(args.tuned) if args.tuned else None
    if tuning is None and fixed is None and args.workload == "c3" and not args.autotune:
        tuning = C3_TUNING
    log(f"workload {args.workload}: setup")
    main_res, plan, segs, wl = run_workload(args.workload, args.steps, args.warmup, torch, G, W, fixed=fixed,
                                            tuning=tuning, specialize=not args.no_specialize)
    log(f"workload {args.workload}: tuning {json.dumps(main_res['tuning'])}")
    log(f"workload {args.workload}: {main_res['batched_ev_s']:.4g} ev/s batched, "
        f"{main_res['plain_ev_s']:.4g} ev/s plain")
    kernel_s = main_res["batched_kernel_ms"] / 1e3
    achieved = main_res["algorithmic_bytes"] / kernel_s / 1e9
    # the device's own streaming copy (2 GiB, read + written bytes): the practical ceiling
    copy_gb_s = G.copy_bandwidth(2 << 30, 5)
    log(f"streaming copy: {copy_gb_s:.0f} GB/s; the step moves {achieved:.0f} GB/s")
    extra = {}
    if not args.no_extra:
        del plan, segs
        torch.cuda.empty_cache()
        # the adjoint first: its FP64 instruction stream is the most sensitive to the clock
        # dropping under minutes of sustained load
        log("workload c4 (forward + adjoint): setup")
        extra["c4"] = run_adjoint(max(5, args.steps), 5, torch, G, W)
        log(f"workload c4: {extra['c4']['fwd_adjoint_ev_s']:.4g} ev/s fwd+adjoint, "
            f"{extra['c4']['fwd_adjoint_over_forward']:.2f}x forward")
        torch.cuda.empty_cache()
        for name in ("c1", "c0", "c2", "c3"):
            if name == args.workload and tuning is None and fixed is None:
                continue
            log(f"workload {name}: setup")
            r, p2, s2, _ = run_workload(name, max(3, args.steps // 2), 5, torch, G, W)
            key = name if name != args.workload else f"{name}_autotuned"
            extra[key] = {k: r[k] for k in ("n_events", "batched_ev_s", "plain_ev_s", "batched_vs_plain",
                                             "batched_kernel_ms", "tuning")}
            extra[key]["tune_summary"] = r["tune"].split(" c0:")[0].split(" c1:")[0]
            extra[key]["launch"] = r["launch"]
            extra[key]["hbm_frac"] = r["algorithmic_bytes"] / (r["batched_kernel_ms"] / 1e3) / 1e9 / HBM_PEAK_GBS
            log(f"workload {name}: {r['batched_ev_s']:.4g} ev/s batched, {r['plain_ev_s']:.4g} plain")
            del p2, s2
            torch.cuda.empty_cache()
    if not args.no_extra:
        for name, scale in (("c1", 1.0), ("c2", 0.1)):
            log(f"inverse mapping {name}: setup")
            extra[f"inverse_{name}"] = run_inverse(name, max(3, args.steps // 4), 1, torch, G, W, scale=scale)
            r = extra[f"inverse_{name}"]
            log(f"inverse mapping {name}: {r['pairs_s']:.4g} (event, channel) pairs/s, {r['hbm_frac']:.3f} of HBM")
            torch.cuda.empty_cache()
        log("per-event beams c1: setup")
        extra["c1_event_beams"] = run_event_beams(max(5, args.steps // 2), torch, G, W)
        log(f"per-event beams c1: {extra['c1_event_beams']['events_s']:.4g} ev/s")
        torch.cuda.empty_cache()
        log("adaptive grids: setup")
        extra["grid_c1"] = run_grid(max(3, args.steps // 4), torch, G, W)
        log(f"adaptive grids: {extra['grid_c1']['randoms_s']:.4g} randoms/s, {extra['grid_c1']['hbm_frac']:.3f} of HBM")
        torch.cuda.empty_cache()
        log("Whizard entry point (host arrays): setup")
        extra["whizard_abi_c1"] = run_whizard_abi(1_000_000, 3, G, W)
        log(f"Whizard entry point: {extra['whizard_abi_c1']['events_s']:.4g} events/s including PCIe copies")
    cpu = None
    if not args.no_cpu:
        threads, _, _ = cpu_threads()
        # 2e6 events of the main workload (1e6 of the 5e7-event mix for the single core)
        scale = 2e6 / wl.n_events
        log(f"cpu baseline, {threads} threads")
        cpu = cpu_baseline(args.workload, W, threads, scale, 10.0)
        log("cpu baseline, one thread")
        extra["cpu_single_core"] = cpu_baseline(args.workload, W, 1, scale / 2, 10.0)
        if not args.no_extra:
            for name, scale in (("c1", 0.01), ("c2", 0.002)):
                if f"inverse_{name}" in extra:
                    extra[f"inverse_{name}"]["cpu_baseline"] = cpu_inverse_baseline(name, scale, W)
    steps = args.steps
    counters = measured_counters(args.workload)
    fp64 = None
    if counters:
        tflops = counters["flops"] / kernel_s / 1e12
        fp64 = {"achieved": tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": tflops / FP64_PEAK_TFLOPS,
                "flops_per_step": counters["flops"], "counted_with": counters["tuning"]}
    out = {
        "metric": "events/s @1 GPU (FP64, <=1e-12 rel vs phs_cpu), batched vs plain; fwd+adjoint ev/s",
        "value": main_res["batched_ev_s"],
        "unit": "events/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["wall_s"] / steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded U(0,1) FP64 randoms generated on the device",
        "config": {"workload": f"{args.workload}: {W.DESCRIPTIONS[args.workload]}",
                   "n_events": main_res["n_events"], "segments": main_res["n_segments"],
                   "launch": main_res["launch"], "tuned": main_res["tuning"],
                   "tune_summary": main_res["tune"].split(" c0:")[0].split(" c1:")[0]},
        "plain_ev_s": main_res["plain_ev_s"],
        "batched_vs_plain": main_res["batched_vs_plain"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": counters["traffic"] if counters else None,
                     "traffic_counted_with": counters["tuning"] if counters else None,
                     "kernel": "forward_kernel (every launch of one step)", "kernel_ms": main_res["batched_kernel_ms"],
                     "bytes_per_launch": main_res["algorithmic_bytes"], "fp64": fp64,
                     "copy_ceiling": {"gb_s": copy_gb_s, "frac": achieved / copy_gb_s,
                                      "what": "the device's streaming copy of 2 GiB, one 16-byte word per thread"}},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
