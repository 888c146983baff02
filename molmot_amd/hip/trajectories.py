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
from .plan import HIPMolData


def recoil_kick_scale(mol_data, hbar: float = HBAR_CODATA) -> float:
    """v_recoil * sqrt(1 + 1/3): per-photon velocity variance of absorption
    along +-z plus isotropic emission projected on z."""
    return float(hbar * mol_data.k / mol_data.mass * np.sqrt(1.0 + 1.0 / 3.0))


def traj_euler_device(plan: HIPMolData, z0: torch.Tensor, v0: torch.Tensor, B_gradient: float, dt: float,
                      n_steps: int, z_escape: float = 0.015, save_full: bool = False,
                      kick_scale: float = 0.0, seed: int = 0, stream_id: int = 0, stream=None):
    """Device tensors in/out: (z_final, v_final, n_actual, z_traj | None, v_traj | None)."""
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
    rc = _lib.lib().mm_traj_euler_1d(plan.handle, n, _lib.ptr(z0), _lib.ptr(v0), float(B_gradient),
                                     plan.mass, float(dt), int(n_steps), float(z_escape),
                                     _lib.ptr(zt), _lib.ptr(vt), _lib.ptr(zf), _lib.ptr(vf), _lib.ptr(na),
                                     float(kick_scale), int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF,
                                     int(stream_id) & 0xFFFFFFFF, _lib.stream_handle(stream))
    _lib.check(rc, "mm_traj_euler_1d")
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
    plan = _plan_for(simulator, cuda_data, hbar)
    n_steps = int(t_max / dt)
    dev = torch.device("cuda", torch.cuda.current_device())
    z0 = torch.as_tensor(np.ascontiguousarray(z0_arr, dtype=np.float64)).to(dev)
    v0 = torch.as_tensor(np.ascontiguousarray(v0_arr, dtype=np.float64)).to(dev)
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
