"""Trajectory outcome checks (molmot/mot/analysis.py:290-355).

Host functions over the (t, z, v) results of simulate_trajectory /
simulate_trajectories: the same rules as the reference, so a list produced by
the GPU integrator gives the reference's numbers.  ``capture_fraction_batch``
applies the rule to a whole device ensemble without building per-particle
tuples."""
from __future__ import annotations

from typing import Tuple

import numpy as np


def survived(result: Tuple[np.ndarray, np.ndarray, np.ndarray], r_max: float = 3e-3) -> bool:
    """True if |z| at the last sample is below r_max (analysis.py:290-307)."""
    _, z_arr, _ = result
    return bool(abs(z_arr[-1]) < r_max)


def capture_fraction(results: list, r_max: float = 3e-3, t_min: float = 1e-3) -> float:
    """Fraction of trajectories that ran at least t_min and have |z| < r_max at
    the first sample with t >= t_min (the last one if none), among those that
    ran at least t_min (analysis.py:314-355)."""
    if not results:
        return 0.0
    n_trapped = n_valid = 0
    for t_arr, z_arr, _ in results:
        if t_arr[-1] < t_min:
            continue
        n_valid += 1
        idx = int(np.searchsorted(t_arr, t_min))
        if idx >= len(t_arr):
            idx = len(t_arr) - 1
        if abs(z_arr[idx]) < r_max:
            n_trapped += 1
    return n_trapped / n_valid if n_valid else 0.0


def capture_fraction_batch(t_arr: np.ndarray, z_traj: np.ndarray, r_max: float = 3e-3,
                           t_min: float = 1e-3) -> float:
    """capture_fraction of an ensemble stored as one time axis t_arr
    (n_steps,) and escape-filled positions z_traj (n, n_steps), as
    simulate_trajectories(save_full=True) returns them."""
    t_arr = np.asarray(t_arr)
    z_traj = np.asarray(z_traj)
    if z_traj.shape[0] == 0 or t_arr[-1] < t_min:
        return 0.0
    idx = min(int(np.searchsorted(t_arr, t_min)), len(t_arr) - 1)
    return int(np.count_nonzero(np.abs(z_traj[:, idx]) < r_max)) / z_traj.shape[0]


def capture_fraction_ensemble(simulator, z0_arr, v0_arr, t_max: float, dt: float = 1e-6, r_max: float = 3e-3,
                              t_min: float = 1e-3, z_escape: float = 0.015) -> float:
    """capture_fraction of simulate_trajectory(simulator, z0, v0, t_max, dt,
    z_escape) over the ensemble, from one launch that stops at the sample
    the rule reads (no trajectories stored): with escape fill, z at that
    sample is the final state of a run of that many samples."""
    import torch
    from ..hip.trajectories import traj_simulator_device
    n_steps = int(t_max / dt)
    if n_steps < 1:
        return 0.0
    t_arr = np.arange(n_steps) * dt
    if t_arr[-1] < t_min or len(z0_arr) == 0:
        return 0.0
    idx = min(int(np.searchsorted(t_arr, t_min)), n_steps - 1)
    dev = torch.device("cuda", torch.cuda.current_device())
    z0 = torch.as_tensor(np.ascontiguousarray(z0_arr, dtype=np.float64)).to(dev)
    v0 = torch.as_tensor(np.ascontiguousarray(v0_arr, dtype=np.float64)).to(dev)
    zf, _, _, _, _ = traj_simulator_device(simulator, z0, v0, dt, idx + 1, z_escape)
    return int((zf.abs() < r_max).sum().item()) / zf.numel()
