"""Differentiable rate-equation force (no reference counterpart: the reference
only takes two-point finite differences, molmot/mot/force_scan.py:84-128, and
scans parameter grids, molmot/mot/force_scan.py:191-301).

Parameter vector theta (float64, 7 entries):
    theta[0:4]  detuning offset of frequency components 0..3 (Hz)
    theta[4]    polarisation split (Hz): beam b moves by split_sign_b * split
    theta[5]    s0 scale (every beam's s0 is multiplied by it)
    theta[6]    B' (T/m)
Beam b of a ``ForceProblem`` has frequency
    freq0_b + theta[comp_b] + split_sign_b * theta[4]
and saturation s0_b * theta[5].  ``rate_force(theta, v, z, problem)`` is a
torch.autograd.Function: its backward is the gfx950 adjoint kernel
(mm_rate_eq_grad_points), which returns dF/dtheta, dF/dz and dF/dv per point.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from .constants import hbar as HBAR_CODATA
from .hip import _lib
from .hip.plan import FORWARD, GRADIENT, HIPMolData
from .obe.fields import BeamTable, LaserBeam, beam_table

N_THETA = 7
J_STRIDE = 9   # 4 detunings, split, s0 scale, B', z, v


@dataclass
class ForceProblem:
    mol_data: object
    beams: List[LaserBeam]
    Gamma_eff_factor: float = 1.0
    hbar: float = HBAR_CODATA

    def theta0(self, B_gradient: float) -> torch.Tensor:
        return torch.tensor([0.0, 0.0, 0.0, 0.0, 0.0, 1.0, float(B_gradient)], dtype=torch.float64)

    def table(self, theta) -> BeamTable:
        th = [float(x) for x in theta]
        bt = beam_table(self.beams)
        if bt.n_comp > 4:
            raise ValueError("at most 4 frequency components carry detuning gradients")
        freqs = bt.freqs + np.array([th[c] for c in bt.comp]) + bt.split_sign * th[4]
        s0s = bt.s0s * th[5]
        return BeamTable(bt.dirs, freqs, bt.pols, s0s, bt.comp, bt.split_sign, bt.n_comp)

    def plan(self, theta) -> HIPMolData:
        return HIPMolData(self.mol_data, self.table(theta), self.Gamma_eff_factor, hbar=self.hbar)


def force_and_jacobian(problem: ForceProblem, theta, v: torch.Tensor, z: torch.Tensor, stream=None,
                       specialize: bool = False, plan: Optional[HIPMolData] = None):
    """F (n,) and J (n, 9) = dF/d[det0..det3, split, s0, B', z, v] on the device.
    ``specialize``: run the plan-specialised gradient kernel (built once per
    set of tables and cached; worth it for repeated maps of one theta).
    ``plan``: problem.plan(theta) built by the caller, to reuse across calls
    at the same theta (otherwise each call builds it)."""
    plan = problem.plan(theta) if plan is None else plan
    if specialize and plan.specializable(GRADIENT):
        plan.specialize(GRADIENT)
    B = float(theta[6])
    v = v.contiguous().reshape(-1)
    z = z.contiguous().reshape(-1)
    n = v.numel()
    F = torch.empty(n, dtype=torch.float64, device=v.device)
    J = torch.empty((n, J_STRIDE), dtype=torch.float64, device=v.device)
    rc = _lib.lib().mm_rate_eq_grad_points(plan.handle, n, _lib.ptr(v), _lib.ptr(z), B, _lib.ptr(F),
                                           _lib.ptr(J), _lib.stream_handle(stream))
    _lib.check(rc, "mm_rate_eq_grad_points")
    return F, J


class _RateForce(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, v, z, problem):
        th = theta.detach().cpu().numpy()
        F, J = force_and_jacobian(problem, th, v.detach(), z.detach())
        ctx.save_for_backward(J)
        ctx.shape = v.shape
        return F.reshape(v.shape)

    @staticmethod
    def backward(ctx, gF):
        (J,) = ctx.saved_tensors
        g = gF.reshape(-1, 1).to(J.dtype)
        gtheta = (g * J[:, :N_THETA]).sum(0).to("cpu")
        gz = (g[:, 0] * J[:, 7]).reshape(ctx.shape)
        gv = (g[:, 0] * J[:, 8]).reshape(ctx.shape)
        return gtheta, gv, gz, None


def rate_force(theta: torch.Tensor, v: torch.Tensor, z: torch.Tensor, problem: ForceProblem) -> torch.Tensor:
    """F(v, z; theta) on the device, differentiable in theta (CPU tensor), v and z."""
    return _RateForce.apply(theta, v, z, problem)


def spring_constant_t(theta, problem: ForceProblem, z0: float = 0.0, dz: float = 0.3e-3, device=None):
    """k = -(F(z0+dz) - F(z0-dz)) / (2 dz) at v = 0 (molmot/mot/force_scan.py:84-104), differentiable."""
    dev = device or torch.device("cuda", torch.cuda.current_device())
    z = torch.tensor([z0 + dz, z0 - dz], dtype=torch.float64, device=dev)
    F = rate_force(theta, torch.zeros_like(z), z, problem)
    return -(F[0] - F[1]) / (2.0 * dz)


def damping_coefficient_t(theta, problem: ForceProblem, z0: float = 0.0, dv: float = 0.1, device=None):
    """alpha = -(F(dv) - F(-dv)) / (2 dv) at z0 (molmot/mot/force_scan.py:107-128), differentiable."""
    dev = device or torch.device("cuda", torch.cuda.current_device())
    v = torch.tensor([dv, -dv], dtype=torch.float64, device=dev)
    F = rate_force(theta, v, torch.full_like(v, z0), problem)
    return -(F[0] - F[1]) / (2.0 * dv)


# ---------------------------------------------------------------------------
# trajectories
# ---------------------------------------------------------------------------

def trajectory_vjp(problem: ForceProblem, theta, z0: torch.Tensor, v0: torch.Tensor, mass: float, dt: float,
                   n_updates: int, cot_z: Optional[torch.Tensor] = None, cot_v: Optional[torch.Tensor] = None,
                   z_escape: float = 0.015, n_up: Optional[torch.Tensor] = None, ckpt_every: int = 500,
                   stream=None, plan: Optional[HIPMolData] = None):
    """Vector-Jacobian product through ``n_updates`` symplectic-Euler steps
    of the deterministic (kick-free) trajectory kernel, for every particle:
    L_i = cot_z[i] z_N + cot_v[i] v_N.  Returns dict(z_final, v_final,
    n_done, g_theta (n, 7), g_z0 (n,), g_v0 (n,)); the gfx950 kernels
    recompute each segment of ``ckpt_every`` steps from a checkpoint with the
    force Jacobian stored per step (workspace ~ 8 (9 K + 2 N/K) bytes per
    particle).  ``plan``: problem.plan(theta) built by the caller, to reuse
    across calls at the same theta (otherwise each call builds it)."""
    plan = problem.plan(theta) if plan is None else plan
    B = float(theta[6])
    dev = z0.device
    z0 = z0.contiguous().reshape(-1).to(torch.float64)
    v0 = v0.contiguous().reshape(-1).to(torch.float64)
    n = z0.numel()
    n_steps = int(n_updates) + 1
    # the forward pass and the segments (one gradient per update) may run
    # plan-specialised kernels (the "auto" rule of HIPMolData)
    plan.maybe_specialize(float(n) * float(n_updates), FORWARD)
    plan.maybe_specialize(float(n) * float(n_updates), GRADIENT)
    L = _lib.lib()
    nw = L.mm_traj_vjp_workspace_doubles(n, n_steps, int(ckpt_every))
    if nw < 0:
        raise ValueError("invalid trajectory VJP sizes")
    work = torch.empty(max(nw, 1), dtype=torch.float64, device=dev)
    f64 = dict(dtype=torch.float64, device=dev)
    out = {"z_final": torch.empty(n, **f64), "v_final": torch.empty(n, **f64),
           "n_done": torch.empty(n, dtype=torch.int32, device=dev), "g_theta": torch.empty((n, N_THETA), **f64),
           "g_z0": torch.empty(n, **f64), "g_v0": torch.empty(n, **f64)}
    cz = None if cot_z is None else cot_z.contiguous().reshape(-1).to(torch.float64)
    cv = None if cot_v is None else cot_v.contiguous().reshape(-1).to(torch.float64)
    nu = None if n_up is None else n_up.contiguous().reshape(-1).to(torch.int32)
    rc = L.mm_traj_vjp_1d(plan.handle, n, _lib.ptr(z0), _lib.ptr(v0), B, float(mass), float(dt), n_steps,
                          _lib.ptr(nu), float(z_escape), int(ckpt_every), _lib.ptr(cz), _lib.ptr(cv),
                          _lib.ptr(out["z_final"]), _lib.ptr(out["v_final"]), _lib.ptr(out["n_done"]),
                          _lib.ptr(out["g_theta"]), _lib.ptr(out["g_z0"]), _lib.ptr(out["g_v0"]), _lib.ptr(work),
                          nw, _lib.stream_handle(stream))
    _lib.check(rc, "mm_traj_vjp_1d")
    return out


def trajectory_tangent(problem: ForceProblem, theta, z0: torch.Tensor, v0: torch.Tensor, mass: float, dt: float,
                       n_updates: int, z_escape: float = 0.015, n_up: Optional[torch.Tensor] = None, stream=None,
                       schedule: str = "auto", check: bool = True, plan: Optional[HIPMolData] = None):
    """Forward sensitivities of the same trajectories in one gfx950 pass:
    dict(z_final, v_final, n_done, sens (n, 2, 9)) with sens[i, 0] =
    d z_N / d p, sens[i, 1] = d v_N / d p for p = (det0..3, split, s0
    scale, B', z_0, v_0).  Any cotangent's VJP is then a contraction
    (``vjp_from_sens``); one force gradient per step.  ``schedule`` as for
    traj_euler_device ("auto", "grid" or "queue"; bit-identical results);
    the queue's small workspace is allocated here.  ``check`` as for
    traj_euler_device: a queue launch is waited for and raises if a queue
    wait gave up.  ``plan``: problem.plan(theta) built by the caller, to reuse
    across calls at the same theta (otherwise each call builds it)."""
    from .hip.trajectories import _SCHEDULES
    if schedule not in _SCHEDULES:
        raise ValueError(f"schedule must be one of {sorted(_SCHEDULES)}")
    plan = problem.plan(theta) if plan is None else plan
    B = float(theta[6])
    dev = z0.device
    z0 = z0.contiguous().reshape(-1).to(torch.float64)
    v0 = v0.contiguous().reshape(-1).to(torch.float64)
    n = z0.numel()
    f64 = dict(dtype=torch.float64, device=dev)
    out = {"z_final": torch.empty(n, **f64), "v_final": torch.empty(n, **f64),
           "n_done": torch.empty(n, dtype=torch.int32, device=dev), "sens": torch.empty((n, 2, N_THETA + 2), **f64)}
    nu = None if n_up is None else n_up.contiguous().reshape(-1).to(torch.int32)
    plan.maybe_specialize(float(n) * float(n_updates), GRADIENT)
    L = _lib.lib()
    ws_bytes = 0 if schedule == "grid" else int(L.mm_traj_euler_workspace_bytes(n, int(n_updates) + 1))
    ws = _lib.queue_workspace(ws_bytes, dev, stream)
    rc = L.mm_traj_tangent_1d_ws(plan.handle, n, _lib.ptr(z0), _lib.ptr(v0), B, float(mass), float(dt),
                                 int(n_updates) + 1, _lib.ptr(nu), float(z_escape), _lib.ptr(out["z_final"]),
                                 _lib.ptr(out["v_final"]), _lib.ptr(out["n_done"]), _lib.ptr(out["sens"]),
                                 _SCHEDULES[schedule], _lib.ptr(ws), ws_bytes, _lib.stream_handle(stream))
    _lib.check(rc, "mm_traj_tangent_1d_ws")
    if ws is not None:
        ws.record_stream(stream if stream is not None else torch.cuda.current_stream(dev))
        if check:
            # a queue wait that gave up leaves n_done = -1 and partial sensitivities
            _lib.check_queue(ws, "mm_traj_tangent_1d_ws", stream)
    return out


def vjp_from_sens(sens: torch.Tensor, cot_z: Optional[torch.Tensor], cot_v: Optional[torch.Tensor]):
    """(g_theta (n, 7), g_z0 (n,), g_v0 (n,)) for L_i = cot_z[i] z_N + cot_v[i] v_N."""
    n = sens.shape[0]
    w = torch.zeros((n, 2), dtype=torch.float64, device=sens.device)
    if cot_z is not None:
        w[:, 0] = cot_z.reshape(-1)
    if cot_v is not None:
        w[:, 1] = cot_v.reshape(-1)
    g = (w[:, :, None] * sens).sum(1)
    return g[:, :N_THETA], g[:, N_THETA], g[:, N_THETA + 1]


class _TrajectoryFinal(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, z0, v0, problem, mass, dt, n_updates, z_escape, ckpt_every, mode):
        th = theta.detach().cpu().numpy()
        ctx.mode = mode
        if mode == "tangent":
            # forward and sensitivities in one pass; backward is a contraction
            o = trajectory_tangent(problem, th, z0.detach(), v0.detach(), mass, dt, n_updates, z_escape)
            ctx.save_for_backward(o["sens"])
            ctx.shape = z0.shape
        else:
            o = trajectory_vjp(problem, th, z0.detach(), v0.detach(), mass, dt, n_updates, None, None, z_escape,
                               None, ckpt_every)
            ctx.args = (th, z0.detach(), v0.detach(), problem, mass, dt, n_updates, z_escape, ckpt_every)
        return o["z_final"].reshape(z0.shape), o["v_final"].reshape(z0.shape)

    @staticmethod
    def backward(ctx, gz, gv):
        if ctx.mode == "tangent":
            (sens,) = ctx.saved_tensors
            gth, gz0, gv0 = vjp_from_sens(sens, gz, gv)
            shape = ctx.shape
        else:
            th, z0, v0, problem, mass, dt, n_updates, z_escape, ckpt_every = ctx.args
            o = trajectory_vjp(problem, th, z0, v0, mass, dt, n_updates,
                               None if gz is None else gz.reshape(-1), None if gv is None else gv.reshape(-1),
                               z_escape, None, ckpt_every)
            gth, gz0, gv0 = o["g_theta"], o["g_z0"], o["g_v0"]
            shape = z0.shape
        return (gth.sum(0).to("cpu"), gz0.reshape(shape), gv0.reshape(shape), None, None, None, None, None, None,
                None)


def trajectory_final(theta: torch.Tensor, z0: torch.Tensor, v0: torch.Tensor, problem: ForceProblem, mass: float,
                     dt: float, n_updates: int, z_escape: float = 0.015, ckpt_every: int = 500,
                     mode: str = "tangent"):
    """(z_N, v_N) after n_updates Euler steps (or at escape), differentiable
    in theta (CPU tensor), z0 and v0 (device tensors).  mode "tangent"
    computes the sensitivities during the forward pass (one force gradient
    per step, no workspace); "checkpoint" runs the plain forward and, in
    backward, the checkpointed adjoint (every ckpt_every steps)."""
    if mode not in ("tangent", "checkpoint"):
        raise ValueError("mode is 'tangent' or 'checkpoint'")
    return _TrajectoryFinal.apply(theta, z0, v0, problem, mass, dt, n_updates, z_escape, ckpt_every, mode)


def capture_fraction_t(theta: torch.Tensor, z0: torch.Tensor, v0: torch.Tensor, problem: ForceProblem,
                       mass: float, dt: float, n_updates: int, r_max: float = 3e-3, width: float = 1e-4,
                       z_escape: float = 0.015) -> torch.Tensor:
    """Smoothed capture fraction, differentiable in theta, z0 and v0:
    mean over particles of sigmoid((r_max - |z_N|) / width), z_N the state
    after n_updates steps (an escaped particle keeps its escape state, as in
    molmot/propagation/trajectories.py:66-70).  As width -> 0 this is the
    rule of capture_fraction (molmot/mot/analysis.py:314-355) at t_min =
    n_updates dt; its gradient comes from the forward sensitivities of every
    trajectory (escape steps held fixed)."""
    zf, _ = trajectory_final(theta, z0, v0, problem, mass, dt, n_updates, z_escape)
    return torch.sigmoid((r_max - zf.abs()) / width).mean()


def _escapes(problem, theta, z_start, vs, mass, dt, n_updates, z_escape):
    from .hip.trajectories import traj_euler_device
    plan = problem.plan(theta)
    dev = torch.device("cuda", torch.cuda.current_device())
    v0 = -torch.as_tensor(vs, dtype=torch.float64, device=dev)
    z0 = torch.full_like(v0, float(z_start))
    if abs(plan.mass - mass) > 1e-12 * mass:
        raise ValueError("mass differs from the molecule's")
    _, _, na, _, _ = traj_euler_device(plan, z0, v0, float(theta[6]), dt, n_updates + 1, z_escape)
    return na.cpu().numpy() < n_updates + 1


def capture_threshold(problem: ForceProblem, theta, z_start: float = 3e-3, v_max: float = 15.0,
                      dt: float = 1e-6, n_updates: int = 5000, z_escape: float = 0.015, tol: float = 1e-9,
                      n_probe: int = 256):
    """Continuous capture velocity v*: the smallest initial speed (from
    z_start towards -z) whose trajectory leaves |z| < z_escape within
    n_updates steps; the reference's capture_velocity
    (molmot/mot/force_scan.py:131-188) is v* rounded down to its dv grid.
    Batched bracketing: n_probe speeds per launch."""
    mass = problem.mol_data.mass
    lo, hi = 0.0, float(v_max)
    if not _escapes(problem, theta, z_start, [hi], mass, dt, n_updates, z_escape)[0]:
        return float("inf")
    while hi - lo > tol:
        vs = np.linspace(lo, hi, n_probe + 2)[1:-1]
        esc = _escapes(problem, theta, z_start, vs, mass, dt, n_updates, z_escape)
        j = int(np.argmax(esc)) if esc.any() else len(vs)
        lo = vs[j - 1] if j > 0 else lo
        hi = vs[j] if j < len(vs) else hi
    return 0.5 * (lo + hi)


def capture_velocity_grad(problem: ForceProblem, theta, z_start: float = 3e-3, v_max: float = 15.0,
                          dt: float = 1e-6, n_updates: int = 5000, z_escape: float = 0.015, ckpt_every: int = 500,
                          mode: str = "tangent"):
    """(v*, dv*/dtheta (7,)) by implicit differentiation at the capture
    threshold.  At v* the far-side turning point z_turn (first step where the
    velocity reaches 0) sits at -z_escape; with z_turn = z_{k*} and v_{k*} ~ 0,
    dv*/dtheta = -(dz_turn/dtheta) / (dz_turn/dv*), both sensitivities from
    one trajectory stopped at k*: forward sensitivities (mode "tangent") or
    the checkpointed adjoint (mode "checkpoint", every ckpt_every steps)."""
    from .hip.trajectories import traj_euler_device
    vstar = capture_threshold(problem, theta, z_start, v_max, dt, n_updates, z_escape)
    if not np.isfinite(vstar):
        raise ValueError("no escape below v_max: the capture threshold is not bracketed")
    vb = vstar * (1.0 - 1e-9)      # just captured: the turning point exists
    plan = problem.plan(theta)
    dev = torch.device("cuda", torch.cuda.current_device())
    z0 = torch.tensor([float(z_start)], dtype=torch.float64, device=dev)
    v0 = torch.tensor([-vb], dtype=torch.float64, device=dev)
    _, _, _, zt, vt = traj_euler_device(plan, z0, v0, float(theta[6]), dt, n_updates + 1, 1e30, save_full=True)
    k = int(torch.argmin(zt[0]).item())      # the far-side turning point
    if k == 0:
        raise ValueError("no turning point within n_updates")
    mass = problem.mol_data.mass
    if mode == "tangent":
        sens = trajectory_tangent(problem, theta, z0, v0, mass, dt, k, 1e30)["sens"][0, 0].cpu().numpy()
        dz_dtheta, dz_dv0 = sens[:N_THETA], float(sens[N_THETA + 1])
    elif mode == "checkpoint":
        o = trajectory_vjp(problem, theta, z0, v0, mass, dt, k, torch.ones(1, dtype=torch.float64, device=dev),
                           None, 1e30, None, ckpt_every)
        dz_dtheta, dz_dv0 = o["g_theta"][0].cpu().numpy(), float(o["g_v0"][0])
    else:
        raise ValueError("mode is 'tangent' or 'checkpoint'")
    dz_dvstar = -dz_dv0          # v_init = -v*
    return vstar, -dz_dtheta / dz_dvstar
