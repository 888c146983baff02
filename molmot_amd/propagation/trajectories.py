"""Trajectories and particle sampling with the reference's signatures
(molmot/propagation/trajectories.py:19-442), integrated on MI355X.

The Euler loops of the reference run as persistent gfx950 kernels (one
launch per call); the samplers are host-side numpy, as in the reference.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np
import torch

from ..constants import k_B
from ..hip import _lib
from ..hip.trajectories import simulate_trajectories_hip, traj_euler_device


def _plan(simulator):
    plan = getattr(simulator, "_hip_plan", None)
    if plan is None:
        raise TypeError("simulator must be a molmot_amd DCMOTSimulator (it owns the device plan)")
    return plan


def _plan_1d(simulator):
    """The plan for the 1D trajectory kernels: None for an RF simulator,
    whose two switching phases the kernel integrates together."""
    return None if hasattr(simulator, "_phase_plans") else _plan(simulator)


def simulate_trajectory(simulator, z0: float, v0: float, t_max: float, dt: float = 1e-6,
                        z_escape: float = 0.015) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(t, z, v) of one particle, Euler (molmot/propagation/trajectories.py:19-75);
    DC or RF simulator."""
    _, _, _, zt, vt = simulate_trajectories_hip(simulator, [z0], [v0], t_max, dt, z_escape, True,
                                                cuda_data=_plan_1d(simulator))
    n_steps = int(t_max / dt)
    return np.arange(n_steps) * dt, zt[0], vt[0]


def simulate_trajectories(simulator, z0_arr, v0_arr, t_max: float, dt: float = 1e-6,
                          z_escape: float = 0.015, save_full: bool = False, recoil: bool = False,
                          seed: int = 0):
    """Ensemble version: (z_final, v_final, n_actual, z_traj | None, v_traj | None)."""
    return simulate_trajectories_hip(simulator, z0_arr, v0_arr, t_max, dt, z_escape, save_full,
                                     cuda_data=_plan_1d(simulator), recoil=recoil, seed=seed)


def simulate_trajectories_3d(simulator, r0_list, v0_list, t_max: float, dt: float = 1e-6,
                             r_escape: float = 0.015) -> List[Tuple[np.ndarray, np.ndarray, np.ndarray]]:
    """Per-axis 3D model (molmot/propagation/trajectories.py:78-156): list of
    (t, r (n, 3), v (n, 3)), one per particle, all particles in one launch."""
    rf = hasattr(simulator, "_phase_plans")
    plans = simulator._phase_plans() if rf else (_plan(simulator),)
    plan = plans[0]
    n_steps = int(t_max / dt)
    r0 = np.ascontiguousarray(np.reshape(np.asarray(r0_list, dtype=np.float64), (-1, 3)))
    v0 = np.ascontiguousarray(np.reshape(np.asarray(v0_list, dtype=np.float64), (-1, 3)))
    n = r0.shape[0]
    dev = torch.device("cuda", torch.cuda.current_device())
    r0t = torch.from_numpy(r0).to(dev)
    v0t = torch.from_numpy(v0).to(dev)
    rfin = torch.empty((n, 3), dtype=torch.float64, device=dev)
    vf = torch.empty((n, 3), dtype=torch.float64, device=dev)
    na = torch.empty(n, dtype=torch.int32, device=dev)
    rt = torch.empty((n, n_steps, 3), dtype=torch.float64, device=dev)
    vt = torch.empty((n, n_steps, 3), dtype=torch.float64, device=dev)
    args = (n, _lib.ptr(r0t), _lib.ptr(v0t), float(simulator.B_gradient), plan.mass, float(dt), n_steps,
            float(r_escape), _lib.ptr(rt), _lib.ptr(vt), _lib.ptr(rfin), _lib.ptr(vf), _lib.ptr(na),
            _lib.stream_handle())
    if rf:   # RF MOT: both switching phases every step
        rc = _lib.lib().mm_traj_axis3d_rf(plans[0].handle, plans[1].handle, *args)
    else:
        rc = _lib.lib().mm_traj_axis3d(plan.handle, *args)
    _lib.check(rc, "mm_traj_axis3d")
    t = np.arange(n_steps) * dt
    rt, vt = rt.cpu().numpy(), vt.cpu().numpy()
    return [(t, rt[i], vt[i]) for i in range(n)]


def monte_carlo_temperature(simulator, n_particles: int = 100, t_max: float = 0.1, dt: float = 1e-6,
                            z_init_sigma: float = 1e-3, v_init_sigma: float = 1.0,
                            t_equilibrium: float = 0.05, schedule: str = "auto") -> float:
    """T = m <v^2> / k_B over steps after t_equilibrium of an ensemble started
    from default_rng(42) Gaussians (molmot/propagation/trajectories.py:159-229).
    ``schedule`` (no reference counterpart) picks the DC launch: "auto",
    "grid" (one thread per particle) or "queue" (work queue); bit-identical
    results.  Raises MolmotHIPError if a queue wait gave up."""
    from ..hip.trajectories import _SCHEDULES
    if schedule not in _SCHEDULES:
        raise ValueError(f"schedule must be one of {sorted(_SCHEDULES)}")
    rf = hasattr(simulator, "_phase_plans")
    plans = simulator._phase_plans() if rf else (_plan(simulator),)
    mass = simulator.mol_data.mass
    n_steps = int(t_max / dt)
    i_eq = int(t_equilibrium / dt)
    rng = np.random.default_rng(42)
    z0 = np.empty(n_particles)
    v0 = np.empty(n_particles)
    for i in range(n_particles):  # same draw order as the reference loop
        z0[i] = rng.normal(0, z_init_sigma)
        v0[i] = rng.normal(0, v_init_sigma)
    dev = torch.device("cuda", torch.cuda.current_device())
    z0t = torch.from_numpy(z0).to(dev)
    v0t = torch.from_numpy(v0).to(dev)
    vsq = torch.empty(n_particles, dtype=torch.float64, device=dev)
    cnt = torch.empty(n_particles, dtype=torch.int64, device=dev)
    args = (n_particles, _lib.ptr(z0t), _lib.ptr(v0t), float(simulator.B_gradient), mass, float(dt), n_steps, i_eq,
            0.015, _lib.ptr(vsq), _lib.ptr(cnt))
    L = _lib.lib()
    if rf:   # RF MOT: both switching phases every step
        plans[0].maybe_specialize_rf(plans[1], n_particles * n_steps)
        rc = L.mm_traj_vsq_rf_1d(plans[0].handle, plans[1].handle, *args, _lib.stream_handle())
    else:
        # the plan-specialised kernels on a large enough run (the plan's
        # "auto" rule; bit-identical sums)
        from ..hip.plan import FORWARD
        plans[0].maybe_specialize(float(n_particles) * float(n_steps), FORWARD)
        # schedule auto: the work queue where one thread per particle would
        # load the SIMDs unevenly (bit-identical results)
        ws_bytes = 0 if schedule == "grid" else int(L.mm_traj_vsq_workspace_bytes(n_particles, n_steps))
        ws = _lib.queue_workspace(ws_bytes, dev)
        rc = L.mm_traj_vsq_1d_ws(plans[0].handle, *args, _SCHEDULES[schedule], _lib.ptr(ws), ws_bytes,
                                 _lib.stream_handle())
    _lib.check(rc, "mm_traj_vsq_1d")
    if int(cnt.min().item()) < 0:
        raise _lib.MolmotHIPError("mm_traj_vsq_1d_ws: a work-queue wait timed out")
    n_samples = int(cnt.sum().item())
    if n_samples == 0:
        return float("inf")
    return float(mass * (vsq.sum().item() / n_samples) / k_B)


# ---------------------------------------------------------------- samplers
# (molmot/propagation/trajectories.py:236-442: host-side numpy, same draws)

def sample_direction(rng=None):
    rng = rng or np.random.default_rng()
    theta = 2.0 * math.pi * rng.random()
    z = 2.0 * rng.random() - 1.0
    rp = math.sqrt(max(0.0, 1.0 - z * z))
    return np.array([rp * math.cos(theta), rp * math.sin(theta), z])


def sample_maxwell_boltzmann(temperature, mass, n_particles=1, rng=None):
    rng = rng or np.random.default_rng()
    sigma = math.sqrt(k_B * temperature / mass) if temperature > 0 else 0.0
    v = rng.normal(0.0, sigma, size=(n_particles, 3)) if sigma > 0 else np.zeros((n_particles, 3))
    return v[0] if n_particles == 1 else v


def sample_maxwell_boltzmann_speed(temperature, mass, n_particles=1, rng=None):
    v = sample_maxwell_boltzmann(temperature, mass, n_particles, rng)
    return float(np.linalg.norm(v)) if n_particles == 1 else np.linalg.norm(v, axis=1)


def sample_gaussian_position(sigma, n_particles=1, rng=None):
    rng = rng or np.random.default_rng()
    r = rng.normal(0.0, sigma, size=(n_particles, 3))
    return r[0] if n_particles == 1 else r


def sample_uniform_sphere(radius, n_particles=1, rng=None):
    rng = rng or np.random.default_rng()
    out = np.zeros((n_particles, 3))
    for i in range(n_particles):
        while True:
            r = rng.uniform(-radius, radius, size=3)
            if np.linalg.norm(r) <= radius:
                out[i] = r
                break
    return out[0] if n_particles == 1 else out


def make_position_sampler(sigma=1e-3, method="gaussian"):
    rng = np.random.default_rng()
    if method == "gaussian":
        return lambda: sample_gaussian_position(sigma, 1, rng)
    if method == "uniform_sphere":
        return lambda: sample_uniform_sphere(sigma, 1, rng)
    raise ValueError(f"Unknown method: {method}")


def make_velocity_sampler(temperature, mass, method="maxwell_boltzmann"):
    rng = np.random.default_rng()
    if method == "maxwell_boltzmann":
        return lambda: sample_maxwell_boltzmann(temperature, mass, 1, rng)
    if method == "fixed":
        return lambda: np.zeros(3)
    raise ValueError(f"Unknown method: {method}")
