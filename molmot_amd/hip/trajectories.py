"""Batched 1D trajectories on MI355X (one persistent kernel launch).

Mirrors simulate_trajectories_cuda / simulate_trajectory_cuda
(molmot_cuda/trajectories_cuda.py:305-452) and adds optional photon-recoil
heating from a counter-based generator (no reference counterpart).
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from ..constants import hbar as HBAR_CODATA
from . import _lib
from .plan import FORWARD, HIPMolData


def recoil_kick_scale(mol_data, hbar: float = HBAR_CODATA) -> float:
    """v_recoil * sqrt(1 + 1/3): per-photon velocity variance of absorption
    along +-z plus isotropic emission projected on z."""
    return float(hbar * mol_data.k / mol_data.mass * np.sqrt(1.0 + 1.0 / 3.0))


_SCHEDULES = {"auto": 0, "grid": 1, "queue": 2}


def traj_euler_device(plan: HIPMolData, z0: torch.Tensor, v0: torch.Tensor, B_gradient: float, dt: float,
                      n_steps: int, z_escape: float = 0.015, save_full: bool = False,
                      kick_scale: float = 0.0, seed: int = 0, stream_id: int = 0, stream=None,
                      schedule: str = "auto", check: bool = True):
    """Device tensors in/out: (z_final, v_final, n_actual, z_traj | None, v_traj | None).

    schedule: "grid" launches one thread per particle; "queue" runs a
    persistent grid that hands (64-particle group, chunk of steps) units out
    of a work queue so that every SIMD carries the same share; "auto" takes
    the queue where the groups would not spread evenly over the SIMDs.  The
    results are bit-identical.  The queue needs save_full=False and a
    workspace of mm_traj_euler_workspace_bytes (allocated here).  With
    ``check`` (and outside graph capture) a queue launch is waited for and
    raises MolmotHIPError if a queue wait gave up (counts -1); check=False
    leaves the launch asynchronous and the -1 counts to the caller."""
    if schedule not in _SCHEDULES:
        raise ValueError(f"schedule must be one of {sorted(_SCHEDULES)}")
    if z0.shape != v0.shape:
        raise ValueError("z0 and v0 must have the same shape")
    z0 = z0.contiguous().reshape(-1)
    v0 = v0.contiguous().reshape(-1)
    n = z0.numel()
    dev = z0.device
    zf = torch.empty(n, dtype=torch.float64, device=dev)
    vf = torch.empty(n, dtype=torch.float64, device=dev)
    na = torch.empty(n, dtype=torch.int32, device=dev)
    zt = torch.empty((n, n_steps), dtype=torch.float64, device=dev) if save_full else None
    vt = torch.empty((n, n_steps), dtype=torch.float64, device=dev) if save_full else None
    L = _lib.lib()
    plan.maybe_specialize(float(n) * float(n_steps), FORWARD)
    ws_bytes = 0 if (save_full or schedule == "grid") else int(L.mm_traj_euler_workspace_bytes(n, int(n_steps)))
    ws = _lib.queue_workspace(ws_bytes, dev, stream)
    rc = L.mm_traj_euler_1d_ws(plan.handle, n, _lib.ptr(z0), _lib.ptr(v0), float(B_gradient),
                               plan.mass, float(dt), int(n_steps), float(z_escape),
                               _lib.ptr(zt), _lib.ptr(vt), _lib.ptr(zf), _lib.ptr(vf), _lib.ptr(na),
                               float(kick_scale), int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF,
                               int(stream_id) & 0xFFFFFFFF, _SCHEDULES[schedule], _lib.ptr(ws), ws_bytes,
                               _lib.stream_handle(stream))
    _lib.check(rc, "mm_traj_euler_1d_ws")
    if ws is not None:
        # the workspace is freed to torch's caching allocator, which reuses
        # it only for work ordered after this launch on the same stream
        ws.record_stream(stream if stream is not None else torch.cuda.current_stream(dev))
        if check:
            _lib.check_queue(ws, "mm_traj_euler_1d_ws", stream)
    return zf, vf, na, zt, vt


def traj_euler_schedule(plan: HIPMolData, n: int, n_steps: int, save_full: bool = False,
                        kick_scale: float = 0.0) -> str:
    """The schedule traj_euler_device(schedule="auto") runs on the current
    device for this size: "grid" or "queue"."""
    rc = _lib.lib().mm_traj_euler_schedule(plan.handle, int(n), int(n_steps), int(bool(save_full)),
                                           float(kick_scale))
    if rc < 0:
        _lib.check(rc, "mm_traj_euler_schedule")
    return {1: "grid", 2: "queue"}[rc]


def traj_euler_rf_device(plan0: HIPMolData, plan1: HIPMolData, z0: torch.Tensor, v0: torch.Tensor,
                         B_gradient: float, dt: float, n_steps: int, z_escape: float = 0.015,
                         save_full: bool = False, kick_scale: float = 0.0, seed: int = 0, stream_id: int = 0,
                         stream=None):
    """RF MOT trajectories: the force of each step is the mean of plan0 at
    +B_gradient and plan1 at -B_gradient.  Outputs as traj_euler_device."""
    if z0.shape != v0.shape:
        raise ValueError("z0 and v0 must have the same shape")
    z0 = z0.contiguous().reshape(-1)
    v0 = v0.contiguous().reshape(-1)
    n = z0.numel()
    dev = z0.device
    zf = torch.empty(n, dtype=torch.float64, device=dev)
    vf = torch.empty(n, dtype=torch.float64, device=dev)
    na = torch.empty(n, dtype=torch.int32, device=dev)
    zt = torch.empty((n, n_steps), dtype=torch.float64, device=dev) if save_full else None
    vt = torch.empty((n, n_steps), dtype=torch.float64, device=dev) if save_full else None
    plan0.maybe_specialize_rf(plan1, n * n_steps)
    rc = _lib.lib().mm_traj_euler_rf_1d(plan0.handle, plan1.handle, n, _lib.ptr(z0), _lib.ptr(v0),
                                        float(B_gradient), plan0.mass, float(dt), int(n_steps), float(z_escape),
                                        _lib.ptr(zt), _lib.ptr(vt), _lib.ptr(zf), _lib.ptr(vf), _lib.ptr(na),
                                        float(kick_scale), int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF,
                                        int(stream_id) & 0xFFFFFFFF, _lib.stream_handle(stream))
    _lib.check(rc, "mm_traj_euler_rf_1d")
    return zf, vf, na, zt, vt


def _plan_for(simulator, cuda_data, hbar):
    if cuda_data is not None:
        return cuda_data
    cached = getattr(simulator, "_hip_plan", None)
    if cached is not None and (hbar is None or cached.hbar_val == hbar):
        return cached
    return HIPMolData(simulator.mol_data, simulator._beams, getattr(simulator, "Gamma_eff_factor", 1.0),
                      hbar=hbar)


def simulate_trajectories_hip(simulator, z0_arr, v0_arr, t_max: float, dt: float = 1e-6,
                              z_escape: float = 0.015, save_full: bool = False,
                              cuda_data: Optional[HIPMolData] = None, block_size: int = 64,
                              hbar: Optional[float] = None, recoil: bool = False, seed: int = 0):
    """(z_final, v_final, n_actual, z_traj | None, v_traj | None), numpy
    (molmot_cuda/trajectories_cuda.py:305-426).  ``block_size`` is accepted
    for signature compatibility; the kernel uses one wavefront per block."""
    n_steps = int(t_max / dt)
    dev = torch.device("cuda", torch.cuda.current_device())
    z0 = torch.as_tensor(np.ascontiguousarray(z0_arr, dtype=np.float64)).to(dev)
    v0 = torch.as_tensor(np.ascontiguousarray(v0_arr, dtype=np.float64)).to(dev)
    if cuda_data is None and hasattr(simulator, "_phase_plans"):
        # RF MOT: both switching phases every step
        p0, p1 = simulator._phase_plans()
        ks = recoil_kick_scale(simulator.mol_data, p0.hbar_val) if recoil else 0.0
        zf, vf, na, zt, vt = traj_euler_rf_device(p0, p1, z0, v0, simulator.B_gradient, dt, n_steps, z_escape,
                                                  save_full, ks, seed)
    else:
        plan = _plan_for(simulator, cuda_data, hbar)
        ks = recoil_kick_scale(simulator.mol_data, plan.hbar_val) if recoil else 0.0
        zf, vf, na, zt, vt = traj_euler_device(plan, z0, v0, simulator.B_gradient, dt, n_steps, z_escape,
                                               save_full, ks, seed)
    out = [zf.cpu().numpy(), vf.cpu().numpy(), na.cpu().numpy()]
    out += [zt.cpu().numpy(), vt.cpu().numpy()] if save_full else [None, None]
    return tuple(out)


def simulate_trajectory_hip(simulator, z0: float, v0: float, t_max: float, dt: float = 1e-6,
                            z_escape: float = 0.015, cuda_data=None, hbar=None):
    """(t, z, v) of one particle (molmot_cuda/trajectories_cuda.py:429-452)."""
    _, _, _, zt, vt = simulate_trajectories_hip(simulator, [z0], [v0], t_max, dt, z_escape, True,
                                                cuda_data, hbar=hbar)
    n_steps = int(t_max / dt)
    return np.arange(n_steps) * dt, zt[0], vt[0]


simulate_trajectories_cuda = simulate_trajectories_hip
simulate_trajectory_cuda = simulate_trajectory_hip


def traj_simulator_device(simulator, z0: torch.Tensor, v0: torch.Tensor, dt: float, n_samples: int,
                          z_escape: float = 0.015, save_full: bool = False, kick_scale: float = 0.0, seed: int = 0,
                          stream=None):
    """traj_euler_device with the simulator's own plan(s): DC simulators
    use their plan, RF simulators both switching phases (traj_euler_rf_device)."""
    if hasattr(simulator, "_phase_plans"):
        p0, p1 = simulator._phase_plans()
        return traj_euler_rf_device(p0, p1, z0, v0, simulator.B_gradient, dt, n_samples, z_escape, save_full,
                                    kick_scale, seed, stream=stream)
    plan = getattr(simulator, "_hip_plan", None)
    if plan is None:
        raise TypeError("simulator must be a molmot_amd DCMOTSimulator or RFMOTSimulator")
    return traj_euler_device(plan, z0, v0, simulator.B_gradient, dt, n_samples, z_escape, save_full, kick_scale,
                             seed, stream=stream)


def capture_velocity_hip(simulator, z_start: float = 3e-3, v_max: float = 15.0, dv: float = 0.5,
                         dt: float = 1e-6, n_steps: int = 5000, z_escape: float = 0.015) -> float:
    """molmot_cuda.capture_velocity_cuda (molmot_cuda/force_scan_cuda.py:533-669):
    test speeds np.arange(dv, v_max + dv, dv), each started at z_start with
    velocity -v0 and captured if |z| stays within z_escape for n_steps
    updates; returns the last captured speed of the run of captured speeds
    from the lowest one.  All speeds in one launch; RF simulators integrate
    both switching phases on the GPU (the reference falls back to the CPU)."""
    v_test = np.arange(dv, v_max + dv, dv, dtype=np.float64)
    if v_test.size == 0:
        return 0.0
    dev = torch.device("cuda", torch.cuda.current_device())
    v0 = -torch.from_numpy(v_test).to(dev)
    z0 = torch.full_like(v0, float(z_start))
    # n_steps updates = n_steps + 1 samples
    _, _, na, _, _ = traj_simulator_device(simulator, z0, v0, dt, n_steps + 1, z_escape)
    trapped = na.cpu().numpy() == n_steps + 1
    v_capture = 0.0
    for vt, ok in zip(v_test, trapped):
        if not ok:
            break
        v_capture = float(vt)
    return v_capture


capture_velocity_cuda = capture_velocity_hip
