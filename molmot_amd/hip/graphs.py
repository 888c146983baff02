"""HIP-graph capture of fixed launch sequences.

Every hot entry point of the C ABI is asynchronous on the stream it is given,
allocates nothing and never synchronises the host, so a sequence of them can
be recorded once into a HIP graph and replayed: one submission instead of one
per kernel.  That pays where a caller repeats the same sequence on new data
of the same shape -- the checkpointed trajectory adjoint (one forward launch,
one launch per segment and two strided copies), or a chain of force maps and
trajectory batches inside an outer optimisation loop.

Inputs and outputs are static device buffers: write new inputs into them in
place (``tensor.copy_``) between replays and read the outputs after
``replay()``.  The recorded launches hold the buffers' addresses, so the
buffers must stay alive and must not be re-allocated while the graph is used.
Kept out of captured sequences: ``hipMemsetAsync`` (cleared memory is written
by kernels instead) and torch reductions, which may record memset nodes.
"""
from __future__ import annotations

from typing import Callable, Optional

import torch

from . import _lib


class CapturedLaunches:
    """Record ``fn`` -- a callable that issues C-ABI launches on torch's
    current stream and allocates nothing inside -- into a HIP graph.

    ``fn`` is run ``warmup`` times eagerly on a side stream first (first-use
    initialisation stays out of the graph), then once under capture."""

    def __init__(self, fn: Callable[[], None], warmup: int = 1):
        _lib.require_gpu()
        self.fn = fn
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                fn()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            fn()
        torch.cuda.synchronize()

    def replay(self) -> None:
        """Submit the recorded launches on the current stream."""
        self.graph.replay()


class TrajectoryAdjointGraph:
    """The checkpointed trajectory adjoint (mm_traj_vjp_1d) of ``n`` particles
    over ``n_updates`` steps, captured once and replayed for new initial
    states and cotangents of the same size.

    Fill ``z0``, ``v0``, ``cot_z``, ``cot_v`` in place, call ``replay()``,
    then read ``z_final``, ``v_final``, ``n_done``, ``g_theta`` (n, 7),
    ``g_z0``, ``g_v0``.  The plan (beams, detunings, B') is fixed at capture."""

    def __init__(self, problem, theta, n: int, mass: float, dt: float, n_updates: int, z_escape: float = 0.015,
                 ckpt_every: int = 500, device: Optional[torch.device] = None):
        dev = device or torch.device("cuda", torch.cuda.current_device())
        self.plan = problem.plan(theta)
        self.B = float(theta[6])
        f64 = dict(dtype=torch.float64, device=dev)
        self.n, self.n_steps, self.K = int(n), int(n_updates) + 1, int(ckpt_every)
        self.mass, self.dt, self.z_escape = float(mass), float(dt), float(z_escape)
        self.z0, self.v0 = torch.zeros(n, **f64), torch.zeros(n, **f64)
        self.cot_z, self.cot_v = torch.zeros(n, **f64), torch.zeros(n, **f64)
        self.z_final, self.v_final = torch.empty(n, **f64), torch.empty(n, **f64)
        self.n_done = torch.empty(n, dtype=torch.int32, device=dev)
        self.g_theta = torch.empty((n, 7), **f64)
        self.g_z0, self.g_v0 = torch.empty(n, **f64), torch.empty(n, **f64)
        L = _lib.lib()
        self.nw = L.mm_traj_vjp_workspace_doubles(self.n, self.n_steps, self.K)
        if self.nw < 0:
            raise ValueError("invalid trajectory VJP sizes")
        self.work = torch.empty(max(self.nw, 1), **f64)
        self._cap = CapturedLaunches(self._launch)

    def _launch(self) -> None:
        p = _lib.ptr
        rc = _lib.lib().mm_traj_vjp_1d(self.plan.handle, self.n, p(self.z0), p(self.v0), self.B, self.mass, self.dt,
                                       self.n_steps, None, self.z_escape, self.K, p(self.cot_z), p(self.cot_v),
                                       p(self.z_final), p(self.v_final), p(self.n_done), p(self.g_theta),
                                       p(self.g_z0), p(self.g_v0), p(self.work), self.nw, _lib.stream_handle())
        _lib.check(rc, "mm_traj_vjp_1d")

    def replay(self) -> None:
        self._cap.replay()
