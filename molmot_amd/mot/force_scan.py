"""Force profiles and MOT figures of merit (molmot/mot/force_scan.py:18-301),
each computed with one batched kernel launch instead of a Python loop."""
from __future__ import annotations

from typing import Dict

import numpy as np
import torch

from ..hip.trajectories import traj_euler_device, traj_simulator_device  # noqa: F401


def force_vs_z(simulator, z_arr, v: float = 0.0):
    """(F (nz,), pop (nz, n_g), R (nz,)) at fixed v (force_scan.py:18-53)."""
    z_arr = np.asarray(z_arr, dtype=np.float64)
    return simulator.force_batch(np.full_like(z_arr, v), z_arr)


def force_vs_v(simulator, v_arr, z: float = 0.0):
    """F (nv,) at fixed z (force_scan.py:56-81)."""
    v_arr = np.asarray(v_arr, dtype=np.float64)
    return simulator.force_batch(v_arr, np.full_like(v_arr, z))[0]


def force_map_2d(simulator, v_arr, z_arr):
    """F, pop, R over the (v, z) grid, v-major (molmot_cuda/rate_equations_cuda.py:640-677)."""
    V, Z = np.meshgrid(np.asarray(v_arr, dtype=np.float64), np.asarray(z_arr, dtype=np.float64),
                       indexing="ij")
    return simulator.force_batch(V, Z)


def spring_constant(simulator, z0: float = 0.0, dz: float = 0.3e-3) -> float:
    """k = -(F(z0+dz) - F(z0-dz)) / (2 dz) at v = 0 (force_scan.py:84-104)."""
    F = simulator.force_batch(np.zeros(2), np.array([z0 + dz, z0 - dz]))[0]
    return float(-(F[0] - F[1]) / (2.0 * dz))


def damping_coefficient(simulator, z0: float = 0.0, dv: float = 0.1) -> float:
    """alpha = -(F(dv) - F(-dv)) / (2 dv) at z0, in N s/m (force_scan.py:107-128)."""
    F = simulator.force_batch(np.array([dv, -dv]), np.full(2, z0))[0]
    return float(-(F[0] - F[1]) / (2.0 * dv))


def capture_velocity(simulator, z_start: float = 3e-3, v_max: float = 15.0, dv: float = 0.5,
                     dt: float = 1e-6, n_steps: int = 5000, z_escape: float = 0.015) -> float:
    """Largest v0 of the sequence dv, 2dv, ... (built by repeated addition as
    in force_scan.py:166-186) such that it and every smaller test velocity
    stay within z_escape for n_steps Euler steps from (z_start, -v0).  All
    test velocities run in one trajectory launch."""
    tests = []
    v_test = dv
    while v_test <= v_max:
        tests.append(v_test)
        v_test += dv
    if not tests:
        return 0.0
    dev = torch.device("cuda", torch.cuda.current_device())
    v0 = -torch.tensor(tests, dtype=torch.float64, device=dev)
    z0 = torch.full_like(v0, float(z_start))
    # the kernel performs n-1 updates for n samples; the reference loop does n_steps
    _, _, na, _, _ = traj_simulator_device(simulator, z0, v0, dt, n_steps + 1, z_escape)
    trapped = (na.cpu().numpy() == n_steps + 1)
    v_capture = 0.0
    for vt, ok in zip(tests, trapped):
        if not ok:
            break
        v_capture = vt
    return v_capture


def optimize_parameters(mol_data, param_ranges: Dict[str, np.ndarray], B_gradient: float,
                        mot_type: str = "dc", s0: float = 1.0, dz: float = 0.3e-3, dv: float = 0.1,
                        verbose: bool = False) -> Dict:
    """Grid scan of spring constant and damping (force_scan.py:191-301)."""
    from .simulator import RFMOTSimulator
    from ..hip.rate_equations import optimize_parameters_hip
    delta_scan = np.asarray(param_ranges["delta"], dtype=float)
    if mot_type == "dc":
        split_scan = np.asarray(param_ranges.get("split", np.array([0.0])), dtype=float)
        return optimize_parameters_hip(mol_data, delta_scan, split_scan, s0, B_gradient, dz, dv,
                                       verbose=verbose)
    if mot_type == "rf":
        mass = mol_data.mass
        K = np.zeros(delta_scan.size)
        beta = np.zeros(delta_scan.size)
        for i, det in enumerate(delta_scan):
            sim = RFMOTSimulator(mol_data, delta_Gamma=det, s0=s0, B_gradient=B_gradient)
            K[i] = spring_constant(sim, dz=dz)
            beta[i] = damping_coefficient(sim, dv=dv) / mass
        ib = int(np.argmax(K * (K > 0) * (beta > 0)))
        return {"delta_scan": delta_scan, "K_arr": K, "beta_arr": beta,
                "delta_best": float(delta_scan[ib]), "k_best": float(K[ib]), "beta_best": float(beta[ib])}
    raise ValueError(f"Unknown mot_type: {mot_type}")
