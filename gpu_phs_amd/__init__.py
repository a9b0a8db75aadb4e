"""gpu_phs for MI355X: recursive phase-space generation (msq mapping, angles, boost chain)
as hand-written HIP kernels behind the reference's C ABI.

The native library ``libgpu_phs.so`` (built from ``csrc/``) is the product; this module binds
its C ABI (include/gpu_phs.h) with ctypes.  PyTorch only provides device memory and streams.
There is no CPU fallback: without the native library every call raises.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from .process import Process  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
# GPU_PHS_LIB selects another build of the library (e.g. the sanitizer build of csrc/Makefile)
LIB_PATH = os.environ.get("GPU_PHS_LIB") or os.path.join(HERE, "libgpu_phs.so")
CSRC = os.path.join(HERE, "csrc")

GPU_PHS_OK = 0


class GpuPhsError(RuntimeError):
    pass


def build(force: bool = True) -> str:
    """Compile libgpu_phs.so for gfx950 (hipcc; works without a GPU)."""
    args = ["make", "-C", CSRC, "-j8"]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True, stdout=subprocess.DEVNULL)
    return LIB_PATH


class _Segment(ctypes.Structure):
    _fields_ = [("process", ctypes.c_int), ("channel", ctypes.c_int), ("n_events", ctypes.c_int64),
                ("x", ctypes.c_void_p), ("p", ctypes.c_void_p), ("factors", ctypes.c_void_p),
                ("volumes", ctypes.c_void_p), ("oks", ctypes.c_void_p)]


_lib = None

_SIGS = {
    "gpu_phs_last_error": (ctypes.c_int, []),
    "gpu_phs_last_error_message": (ctypes.c_char_p, []),
    "gpu_phs_process_create": (ctypes.c_int, [ctypes.c_int] * 3 + [ctypes.c_void_p] * 9 + [ctypes.c_double] * 2),
    "gpu_phs_process_load_ref": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_double, ctypes.c_double]),
    "gpu_phs_process_destroy": (ctypes.c_int, [ctypes.c_int]),
    "gpu_phs_process_info": (ctypes.c_int, [ctypes.c_int] + [ctypes.c_void_p] * 3),
    "gpu_phs_process_trees": (ctypes.c_int, [ctypes.c_int] + [ctypes.c_void_p] * 7),
    "gpu_phs_process_shapes": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_int]),
    "gpu_phs_plan_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "gpu_phs_plan_destroy": (ctypes.c_int, [ctypes.c_int]),
    "gpu_phs_plan_set_config": (ctypes.c_int, [ctypes.c_int] * 7),
    "gpu_phs_plan_get_config": (ctypes.c_int, [ctypes.c_int] + [ctypes.c_void_p] * 6),
    "gpu_phs_plan_run": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p]),
    "gpu_phs_plan_run_plain": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p]),
    "gpu_phs_plan_run_graph": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p]),
    "gpu_phs_plan_autotune": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
    "gpu_phs_plan_info": (ctypes.c_int, [ctypes.c_int] + [ctypes.c_void_p] * 3),
    "gpu_phs_plan_adjoint": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_void_p, ctypes.c_void_p]),
    "gpu_phs_plan_n_params": (ctypes.c_int, [ctypes.c_int]),
    "gpu_phs_plan_param_map": (ctypes.c_int, [ctypes.c_int] + [ctypes.c_void_p] * 3 + [ctypes.c_int]),
    "gpu_phs_plan_inverse": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "gpu_phs_plan_set_event_beams": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "gpu_phs_plan_set_msq_output": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
    "gpu_phs_plan_set_class_config": (ctypes.c_int, [ctypes.c_int] * 7),
    "gpu_phs_plan_get_class_config": (ctypes.c_int, [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 5),
    "gpu_phs_plan_set_schedule": (ctypes.c_int, [ctypes.c_int] * 3),
    "gpu_phs_plan_get_schedule": (ctypes.c_int, [ctypes.c_int] + [ctypes.c_void_p] * 2),
    "gpu_phs_plan_set_merge_below": (ctypes.c_int, [ctypes.c_int, ctypes.c_int64]),
    "gpu_phs_plan_get_merge_below": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p]),
    "gpu_phs_plan_set_streams": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "gpu_phs_copy_bandwidth": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    "gpu_phs_plan_get_streams": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p]),
    "gpu_phs_plan_set_specialize": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "gpu_phs_grid_create": (ctypes.c_int, [ctypes.c_int]),
    "gpu_phs_grid_destroy": (ctypes.c_int, [ctypes.c_int]),
    "gpu_phs_grid_get_weights": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_int]),
    "gpu_phs_grid_set_weights": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_int]),
    "gpu_phs_grid_generate": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]),
    "gpu_phs_grid_update": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]),
}


class _AdjSegment(ctypes.Structure):
    _fields_ = [("g_p", ctypes.c_void_p), ("g_f", ctypes.c_void_p), ("g_x", ctypes.c_void_p),
                ("loss", ctypes.c_void_p)]


def lib():
    """The loaded native library.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GpuPhsError(f"{LIB_PATH} is missing: build it first (gpu_phs_amd.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc, what=""):
    if rc < 0:
        msg = lib().gpu_phs_last_error_message().decode()
        raise GpuPhsError(f"{what}: {msg} ({rc})")
    return rc


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _stream_ptr(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


class ProcessHandle:
    """A compiled process (its trees turned into kernel programs for fixed beams)."""

    def __init__(self, proc: Process, E1=None, E2=None):
        self.proc = proc
        E1 = proc.sqrts / 2 if E1 is None else E1
        E2 = proc.sqrts / 2 if E2 is None else E2
        d1, d2, hc, mid, mm, mw, flv = proc.ctypes_args()
        fr = np.ascontiguousarray(proc.friends_or_zero(), np.int32)
        flw = np.ascontiguousarray(proc.flv_widths, np.float64)
        self._keep = (d1, d2, hc, mid, mm, mw, flv, fr, flw)
        self.handle = _check(lib().gpu_phs_process_create(
            proc.n_in, proc.n_out, proc.n_channels, _ptr(d1), _ptr(d2), _ptr(hc), _ptr(fr), _ptr(mid),
            _ptr(mm), _ptr(mw), _ptr(flv), _ptr(flw), E1, E2), "process_create")

    @property
    def n_x(self):
        return self.proc.n_x

    @property
    def n_out(self):
        return self.proc.n_out

    def shapes(self):
        """Per channel: the tree shape its forward kernel is specialised for, or -1."""
        out = np.zeros(self.proc.n_channels, np.int32)
        _check(lib().gpu_phs_process_shapes(self.handle, _ptr(out), out.size), "process_shapes")
        return out

    def close(self):
        if self.handle >= 0 and _lib is not None:
            lib().gpu_phs_process_destroy(self.handle)
        self.handle = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_bandwidth(nbytes: int = 2 << 30, reps: int = 5, stream=None) -> float:
    """The device's streaming copy bandwidth in GB/s (read + written bytes per second)."""
    v = ctypes.c_double()
    _check(lib().gpu_phs_copy_bandwidth(int(nbytes), int(reps), _stream_ptr(stream), ctypes.byref(v)),
           "copy_bandwidth")
    return v.value


def load_ref(path: str, E1: float, E2: float) -> int:
    """Process handle from a reference-format .ref file (trees only)."""
    return _check(lib().gpu_phs_process_load_ref(path.encode(), E1, E2), "process_load_ref")


class EventBuffers:
    """Device buffers of one segment (torch tensors on the current device)."""

    def __init__(self, x, p, factor, volume, ok):
        self.x, self.p, self.factor, self.volume, self.ok = x, p, factor, volume, ok

    @staticmethod
    def allocate(x, n_out):
        import torch
        n = x.shape[0]
        dev = x.device
        return EventBuffers(x, torch.empty((n, n_out, 4), dtype=torch.float64, device=dev),
                            torch.empty(n, dtype=torch.float64, device=dev),
                            torch.empty(n, dtype=torch.float64, device=dev),
                            torch.empty(n, dtype=torch.uint8, device=dev))


class Plan:
    """A batch of (process, channel, event range) segments run as one launch.

    ``segments``: list of ``(ProcessHandle, channel, EventBuffers)``.  The tensors must stay
    alive and keep their storage while the plan is used.
    """

    def __init__(self, segments, **config):
        arr = (_Segment * len(segments))()
        self._keep = []
        for i, (ph, ch, buf) in enumerate(segments):
            for t in (buf.x, buf.p, buf.factor, buf.volume, buf.ok):
                if not t.is_contiguous():
                    raise ValueError("segment tensors must be contiguous")
            if buf.x.shape[1] != ph.n_x or tuple(buf.p.shape[1:]) != (ph.n_out, 4):
                raise ValueError("segment buffer shapes do not match the process")
            arr[i] = _Segment(ph.handle, ch, buf.x.shape[0], buf.x.data_ptr(), buf.p.data_ptr(),
                              buf.factor.data_ptr(), buf.volume.data_ptr(), buf.ok.data_ptr())
            self._keep.append((ph, buf))
        self.handle = _check(lib().gpu_phs_plan_create(arr, len(segments)), "plan_create")
        if config:
            self.set_config(**config)

    def set_merge_below(self, events):
        """Tree shapes with fewer events than this in a tile class run in the class's generic
        launch (0: every shape its own specialised launch)."""
        _check(lib().gpu_phs_plan_set_merge_below(self.handle, int(events)), "plan_set_merge_below")

    def set_streams(self, n_streams):
        """Spread the plan's launches over 1-4 streams (1: the tile classes as the schedule
        says)."""
        _check(lib().gpu_phs_plan_set_streams(self.handle, int(n_streams)), "plan_set_streams")

    def set_config(self, **config):
        """Launch configuration: block, ept, grid, order, waves_per_eu, lds; keys left out
        keep their current values (the library's defaults for a new plan)."""
        unknown = set(config) - {"block", "ept", "grid", "order", "waves_per_eu", "lds"}
        if unknown:
            raise TypeError(f"unknown launch parameters {sorted(unknown)}")
        c = self.config()
        c.update(config)
        _check(lib().gpu_phs_plan_set_config(self.handle, c["block"], c["ept"], c["grid"], c["order"],
                                             c["waves_per_eu"], c["lds"]), "plan_set_config")

    def config(self):
        v = [ctypes.c_int() for _ in range(6)]
        _check(lib().gpu_phs_plan_get_config(self.handle, *[ctypes.byref(q) for q in v]), "plan_get_config")
        return dict(block=v[0].value, ept=v[1].value, grid=v[2].value, order=v[3].value,
                    waves_per_eu=v[4].value, lds=v[5].value)

    def tuning(self):
        """The launch shape of both tile classes and the schedule, as the autotuner left them:
        dict(classes=[{block, ept, grid, waves_per_eu, lds}] * 2, order, serial, merge_below, streams)."""
        classes = []
        for c in range(2):
            v = [ctypes.c_int() for _ in range(5)]
            _check(lib().gpu_phs_plan_get_class_config(self.handle, c, *[ctypes.byref(q) for q in v]),
                   "plan_get_class_config")
            classes.append(dict(zip(("block", "ept", "grid", "waves_per_eu", "lds"), (q.value for q in v))))
        o, se = ctypes.c_int(), ctypes.c_int()
        _check(lib().gpu_phs_plan_get_schedule(self.handle, ctypes.byref(o), ctypes.byref(se)), "plan_get_schedule")
        mb = ctypes.c_int64()
        _check(lib().gpu_phs_plan_get_merge_below(self.handle, ctypes.byref(mb)), "plan_get_merge_below")
        ns = ctypes.c_int()
        _check(lib().gpu_phs_plan_get_streams(self.handle, ctypes.byref(ns)), "plan_get_streams")
        return dict(classes=classes, order=o.value, serial=se.value, merge_below=mb.value, streams=ns.value)

    def apply_tuning(self, t):
        """Set what ``tuning()`` returned (e.g. from an earlier autotuner run)."""
        for c, lc in enumerate(t["classes"]):
            _check(lib().gpu_phs_plan_set_class_config(self.handle, c, lc["block"], lc["ept"], lc["grid"],
                                                       lc["waves_per_eu"], lc["lds"]), "plan_set_class_config")
        _check(lib().gpu_phs_plan_set_schedule(self.handle, t["order"], t["serial"]), "plan_set_schedule")
        if "merge_below" in t:
            _check(lib().gpu_phs_plan_set_merge_below(self.handle, t["merge_below"]), "plan_set_merge_below")
        if "streams" in t:
            _check(lib().gpu_phs_plan_set_streams(self.handle, t["streams"]), "plan_set_streams")

    def info(self):
        v = [ctypes.c_int() for _ in range(3)]
        _check(lib().gpu_phs_plan_info(self.handle, *[ctypes.byref(q) for q in v]), "plan_info")
        return dict(n_tiles=v[0].value, n_segments=v[1].value, n_banks=v[2].value)

    def set_event_beams(self, segment, E1=None, E2=None):
        """Per-event beam energies (cuda float64 tensors [n_events] of the segment), or None for
        the process's fixed beams.  The tensors must stay alive while the plan is used."""
        if (E1 is None) != (E2 is None):
            raise ValueError("give both E1 and E2, or neither")
        if E1 is not None:
            n = self._keep[segment][1].x.shape[0]
            for t in (E1, E2):
                if t.dtype != self._keep[segment][1].x.dtype or not t.is_contiguous() or t.shape != (n,):
                    raise ValueError("beam energies must be contiguous float64 tensors of the segment's length")
            self._beams = getattr(self, "_beams", {})
            self._beams[segment] = (E1, E2)
        _check(lib().gpu_phs_plan_set_event_beams(self.handle, segment,
                                                  None if E1 is None else ctypes.c_void_p(E1.data_ptr()),
                                                  None if E2 is None else ctypes.c_void_p(E2.data_ptr())),
               "plan_set_event_beams")

    def set_msq_output(self, segment, msq=None):
        """Per-branch msq of every event of the segment into ``msq`` (a cuda float64 tensor
        [n_events][2^n_out - 1], index bincode - 1; slots of other bincodes are left as they
        are), or None to turn it off.  The tensor must stay alive while the plan is used."""
        if msq is not None:
            buf = self._keep[segment][1]
            n_out = buf.p.shape[1]
            if msq.dtype != buf.x.dtype or not msq.is_contiguous() or tuple(msq.shape) != (buf.x.shape[0], 2 ** n_out - 1):
                raise ValueError("msq must be a contiguous float64 tensor [n_events][2^n_out - 1]")
            self._msq = getattr(self, "_msq", {})
            self._msq[segment] = msq
        _check(lib().gpu_phs_plan_set_msq_output(self.handle, segment,
                                                 None if msq is None else ctypes.c_void_p(msq.data_ptr())),
               "plan_set_msq_output")

    def set_specialize(self, on=True):
        """Shape-specialised kernels for trees of up to four final-state particles (default on)."""
        _check(lib().gpu_phs_plan_set_specialize(self.handle, 1 if on else 0), "plan_set_specialize")

    def run(self, stream=None):
        _check(lib().gpu_phs_plan_run(self.handle, _stream_ptr(stream)), "plan_run")

    def run_plain(self, stream=None):
        _check(lib().gpu_phs_plan_run_plain(self.handle, _stream_ptr(stream)), "plan_run_plain")

    def run_graph(self, stream=None):
        _check(lib().gpu_phs_plan_run_graph(self.handle, _stream_ptr(stream)), "plan_run_graph")

    def inverse(self, seg_rows=None, n_rows=None, factors=None, x_out=None, block=0, stream=None):
        """All-channel inverse mapping of the plan's events (after a forward run).  Rows are
        the segments' events in segment order unless ``seg_rows`` gives each segment's first
        row.  Returns (factors [n_rows][n_channels], x_out [n_channels][n_rows][n_x])."""
        import torch
        ns = [buf.x.shape[0] for _, buf in self._keep]
        if seg_rows is None:
            seg_rows = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
        seg_rows = np.ascontiguousarray(seg_rows, np.int64)
        if n_rows is None:
            n_rows = int(max(r + n for r, n in zip(seg_rows, ns)))
        ph = self._keep[0][0]
        n_ch = ph.proc.n_channels
        dev = self._keep[0][1].x.device
        if factors is None:
            factors = torch.empty((n_rows, n_ch), dtype=torch.float64, device=dev)
        if x_out is None:
            x_out = torch.empty((n_ch, n_rows, ph.n_x), dtype=torch.float64, device=dev)
        _check(lib().gpu_phs_plan_inverse(self.handle, _ptr(seg_rows), n_rows, ctypes.c_void_p(factors.data_ptr()),
                                          ctypes.c_void_p(x_out.data_ptr()), block, _stream_ptr(stream)),
               "plan_inverse")
        return factors, x_out

    def n_params(self):
        return _check(lib().gpu_phs_plan_n_params(self.handle), "plan_n_params")

    def param_map(self):
        """(process handle, channel, bincode) of every (M, Gamma) parameter slot."""
        n = self.n_params() // 2
        pr, ch, bc = (np.zeros(n, np.int32) for _ in range(3))
        _check(lib().gpu_phs_plan_param_map(self.handle, _ptr(pr), _ptr(ch), _ptr(bc), n), "plan_param_map")
        return pr, ch, bc

    def adjoint(self, g_x, g_p=None, g_f=None, loss=None, loss_mode=0, write_forward=True, gparam=None,
                stream=None):
        """Reverse-mode pass over every segment: fills the per-segment tensors g_x (and loss), and
        returns gparam [n_params] (d/dM, d/dGamma interleaved per slot of param_map())."""
        import torch
        n = len(self._keep)
        arr = (_AdjSegment * n)()
        keep = []
        for i in range(n):
            def ptr(t):
                if t is None or t[i] is None:
                    return None
                if not t[i].is_contiguous():
                    raise ValueError("adjoint tensors must be contiguous")
                keep.append(t[i])
                return t[i].data_ptr()
            arr[i] = _AdjSegment(ptr(g_p), ptr(g_f), ptr(g_x), ptr(loss))
        if gparam is None:
            dev = self._keep[0][1].x.device
            gparam = torch.empty(max(self.n_params(), 1), dtype=torch.float64, device=dev)
        _check(lib().gpu_phs_plan_adjoint(self.handle, arr, loss_mode, 1 if write_forward else 0,
                                          ctypes.c_void_p(gparam.data_ptr()), _stream_ptr(stream)), "plan_adjoint")
        return gparam

    def autotune(self, reps=3, stream=None):
        buf = ctypes.create_string_buffer(8192)
        _check(lib().gpu_phs_plan_autotune(self.handle, _stream_ptr(stream), reps, buf, len(buf)),
               "plan_autotune")
        return buf.value.decode()

    def close(self):
        if self.handle >= 0 and _lib is not None:
            lib().gpu_phs_plan_destroy(self.handle)
        self.handle = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Grid:
    """Adaptive grids of a process's randoms: per channel and random, 10 equal bins with
    weights (uniform at first), as the reference's rng.cpp keeps them.  ``generate`` fills the
    randoms of every segment of a plan on the device; ``update`` sets the weights to the
    histogram of the accepted events' randoms."""

    BINS = 10

    def __init__(self, ph: ProcessHandle):
        self.ph = ph
        self.handle = _check(lib().gpu_phs_grid_create(ph.handle), "grid_create")
        self.shape = (ph.proc.n_channels, ph.n_x, self.BINS)

    def weights(self):
        w = np.zeros(self.shape)
        _check(lib().gpu_phs_grid_get_weights(self.handle, _ptr(w), w.size), "grid_get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float64)
        if w.shape != self.shape:
            raise ValueError(f"weights must have shape {self.shape}")
        _check(lib().gpu_phs_grid_set_weights(self.handle, _ptr(w), w.size), "grid_set_weights")

    def generate(self, plan, seed, stream=None):
        _check(lib().gpu_phs_grid_generate(self.handle, plan.handle, seed, _stream_ptr(stream)), "grid_generate")

    def update(self, plan, stream=None):
        _check(lib().gpu_phs_grid_update(self.handle, plan.handle, _stream_ptr(stream)), "grid_update")

    def close(self):
        if self.handle > 0 and _lib is not None:
            lib().gpu_phs_grid_destroy(self.handle)
        self.handle = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def generate(ph: ProcessHandle, x, channel: int, stream=None, **cfg):
    """One-channel convenience: forward pass of x (device [n, n_x] float64)."""
    buf = EventBuffers.allocate(x, ph.n_out)
    plan = Plan([(ph, channel, buf)], **cfg)
    plan.run(stream)
    return buf, plan
