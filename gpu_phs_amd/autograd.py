"""torch.autograd bindings of the forward and adjoint kernels.

``PhaseSpace`` wraps one process with fixed channel segments: ``forward(x, M, Gamma)`` returns
the final-state momenta ``p [n, n_out, 4]``, the Jacobian factor ``[n]``, the volume ``[n]``
and ``ok [n]`` for randoms ``x [n, n_x]`` (events ordered by channel, ``counts[c]`` events of
channel ``c``) and map masses / widths ``M, Gamma [n_channels, 2**n_out - 1]`` (index =
bincode - 1, as the reference's Whizard interface takes them).  Gradients flow to ``x``, ``M``
and ``Gamma`` through the hand-written adjoint kernel; non-ok events get zero gradients.
``M`` and ``Gamma`` enter the compiled kernel programs, so a forward with new values
recompiles the process on the host.
"""
from __future__ import annotations

import numpy as np
import torch

from . import EventBuffers, Plan, ProcessHandle
from .process import Process


class _PhaseSpaceFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, M, Gamma, owner):
        plan, bufs = owner._plan_for(x, M, Gamma)
        plan.run()
        p = torch.cat([b.p for b in bufs])
        f = torch.cat([b.factor for b in bufs])
        v = torch.cat([b.volume for b in bufs])
        ok = torch.cat([b.ok for b in bufs]).bool()
        ctx.owner = owner
        ctx.plan = plan
        ctx.bufs = bufs
        ctx.mark_non_differentiable(v, ok)
        return p, f, v, ok

    @staticmethod
    def backward(ctx, gp, gf, gv, gok):
        owner, plan, bufs = ctx.owner, ctx.plan, ctx.bufs
        counts = [b.x.shape[0] for b in bufs]
        gp = (torch.zeros_like(torch.cat([b.p for b in bufs])) if gp is None else gp).contiguous()
        gf = (torch.zeros_like(torch.cat([b.factor for b in bufs])) if gf is None else gf).contiguous()
        gps, gfs = torch.split(gp, counts), torch.split(gf, counts)
        gxs = [torch.empty_like(b.x) for b in bufs]
        gparam = plan.adjoint(gxs, g_p=list(gps), g_f=list(gfs), loss_mode=0, write_forward=False)
        gM, gW = owner._scatter_params(gparam)
        return torch.cat(gxs), gM, gW, None


class PhaseSpace:
    """A process with a fixed split of events over its channels, differentiable in x, M, Gamma."""

    def __init__(self, proc: Process, counts, **launch):
        self.proc = proc
        self.counts = [int(c) for c in counts]
        if len(self.counts) != proc.n_channels:
            raise ValueError("one count per channel")
        self.launch = launch
        self._key = None

    def _plan_for(self, x, M, Gamma):
        key = (x.data_ptr(), tuple(x.shape), M.detach().cpu().numpy().tobytes(),
               Gamma.detach().cpu().numpy().tobytes())
        if key != self._key:
            proc = self.proc.subset(range(self.proc.n_channels))
            proc.map_masses = M.detach().cpu().numpy().astype(np.float64).reshape(proc.map_masses.shape)
            proc.map_widths = Gamma.detach().cpu().numpy().astype(np.float64).reshape(proc.map_widths.shape)
            self._ph = ProcessHandle(proc)
            xs = torch.split(x.detach(), self.counts)
            self._bufs = [EventBuffers.allocate(xi.contiguous(), proc.n_out) for xi in xs]
            segs = [(self._ph, c, b) for c, b in enumerate(self._bufs) if self.counts[c] > 0]
            self._plan = Plan(segs, **self.launch)
            _, ch, bc = self._plan.param_map()
            self._slots = (ch, bc)
            self._key = key
        return self._plan, [b for c, b in enumerate(self._bufs) if self.counts[c] > 0]

    def _scatter_params(self, gparam):
        ch, bc = self._slots
        g = gparam.detach().cpu().numpy()
        gM = np.zeros(self.proc.map_masses.shape)
        gW = np.zeros(self.proc.map_widths.shape)
        gM[ch, bc - 1] = g[0::2][:len(ch)]
        gW[ch, bc - 1] = g[1::2][:len(ch)]
        return torch.from_numpy(gM), torch.from_numpy(gW)

    def __call__(self, x, M=None, Gamma=None):
        if M is None:
            M = torch.from_numpy(self.proc.map_masses.copy())
        if Gamma is None:
            Gamma = torch.from_numpy(self.proc.map_widths.copy())
        return _PhaseSpaceFn.apply(x, M, Gamma, self)
