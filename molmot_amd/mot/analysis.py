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


# ---------------------------------------------------------------------------
# cloud size, temperature and density vs time (molmot/mot/analysis.py:128-288)
# ---------------------------------------------------------------------------

R_ALIVE = 0.015   # analysis.py:176, 232: a molecule counts while |z| < 15 mm


def _bins(t_max: float, time_bins: int):
    t_edges = np.linspace(0, t_max, time_bins + 1)
    return 0.5 * (t_edges[:-1] + t_edges[1:])


def _sample_index(t_arr, t_mid):
    idx = int(np.searchsorted(t_arr, t_mid))
    return min(idx, len(t_arr) - 1)


def sigma_vs_time(results: list, time_bins: int = 50):
    """(t_centres, sigma_z, sigma_v, n_alive) of a list of (t, z, v)
    trajectories (analysis.py:128-186): the sample nearest each time-bin
    centre, molecules with |z| < 15 mm, population standard deviations; nan
    where fewer than two are alive."""
    t_max = min(r[0][-1] for r in results)
    tc = _bins(t_max, time_bins)
    sz, sv, na = np.zeros(time_bins), np.zeros(time_bins), np.zeros(time_bins)
    for k, tm in enumerate(tc):
        z = np.array([zr[_sample_index(t, tm)] for t, zr, _ in results])
        v = np.array([vr[_sample_index(t, tm)] for t, _, vr in results])
        alive = np.abs(z) < R_ALIVE
        na[k] = np.sum(alive)
        if na[k] > 1:
            sz[k], sv[k] = np.std(z[alive]), np.std(v[alive])
        else:
            sz[k] = sv[k] = np.nan
    return tc, sz, sv, na


def temperature_vs_time(results: list, mass: float, time_bins: int = 50):
    """(t_centres, T) with T = m <v^2> / k_B over the molecules with
    |z| < 15 mm at each bin's sample (analysis.py:194-240)."""
    from ..constants import k_B
    t_max = min(r[0][-1] for r in results)
    tc = _bins(t_max, time_bins)
    T = np.zeros(time_bins)
    for k, tm in enumerate(tc):
        vs = [vr[_sample_index(t, tm)] for t, zr, vr in results if abs(zr[_sample_index(t, tm)]) < R_ALIVE]
        T[k] = mass * np.mean(np.array(vs) ** 2) / k_B if len(vs) > 1 else np.nan
    return tc, T


def _density(sz, na, n_results, N_total):
    n0 = np.zeros(len(sz))
    for k, s in enumerate(sz):
        if np.isfinite(s) and s > 0:
            n_eff = N_total * (na[k] / n_results) if n_results > 0 else 0
            n0[k] = n_eff / (s ** 3 * (2 * np.pi) ** 1.5)
    return n0


def density_vs_time(results: list, N_total: int = 2000, time_bins: int = 50):
    """(t_centres, n0): Gaussian peak density N_alive / (sigma_z^3 (2 pi)^1.5)
    with N scaled by the surviving fraction (analysis.py:248-288)."""
    tc, sz, _, na = sigma_vs_time(results, time_bins)
    return tc, _density(sz, na, len(results), N_total)


def ensemble_time_statistics(simulator, z0_arr, v0_arr, t_max: float, dt: float = 1e-6, time_bins: int = 50,
                             z_escape: float = 0.015, kick_scale: float = 0.0, seed: int = 0, stream=None):
    """The per-bin moments the three functions above read, for the ensemble
    simulate_trajectory would integrate from (z0, v0) (DC simulators), on
    the device in one launch without storing trajectories
    (mm_traj_stats_1d).  Returns dict(t_centres, sample_steps, n_alive,
    mean_z, m2_z, mean_v, m2_v) (numpy; m2 = sum of squared deviations)."""
    import torch
    from ..hip import _lib
    plan = getattr(simulator, "_hip_plan", None)
    if plan is None or hasattr(simulator, "_phase_plans"):
        raise TypeError("ensemble_time_statistics needs a molmot_amd DCMOTSimulator")
    n_steps = int(t_max / dt)
    t_arr = np.arange(n_steps) * dt
    tc = _bins(t_arr[-1], time_bins)
    samp = np.array([_sample_index(t_arr, tm) for tm in tc], dtype=np.int32)
    dev = torch.device("cuda", torch.cuda.current_device())
    z0 = torch.as_tensor(np.ascontiguousarray(z0_arr, dtype=np.float64)).to(dev).reshape(-1)
    v0 = torch.as_tensor(np.ascontiguousarray(v0_arr, dtype=np.float64)).to(dev).reshape(-1)
    n = z0.numel()
    L = _lib.lib()
    # the plan-specialised kernels on a large enough run (the plan's "auto"
    # rule; bit-identical moments)
    from ..hip.plan import FORWARD
    plan.maybe_specialize(float(n) * float(n_steps), FORWARD)
    st = torch.from_numpy(samp).to(dev)
    out = torch.empty((time_bins, 5), dtype=torch.float64, device=dev)
    wb = int(L.mm_traj_stats_workspace_bytes(n, time_bins))
    ws = torch.empty(max(wb, 8), dtype=torch.uint8, device=dev)
    rc = L.mm_traj_stats_1d(plan.handle, n, _lib.ptr(z0), _lib.ptr(v0), float(simulator.B_gradient), plan.mass,
                            float(dt), n_steps, float(z_escape), R_ALIVE, _lib.ptr(st), time_bins, _lib.ptr(out),
                            float(kick_scale), int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF, 0,
                            _lib.ptr(ws), wb, _lib.stream_handle(stream))
    _lib.check(rc, "mm_traj_stats_1d")
    o = out.cpu().numpy()
    return dict(t_centres=tc, sample_steps=samp, n_alive=o[:, 0], mean_z=o[:, 1], m2_z=o[:, 2], mean_v=o[:, 3],
                m2_v=o[:, 4], n_particles=n)


def sigma_vs_time_ensemble(simulator, z0_arr, v0_arr, t_max: float, dt: float = 1e-6, time_bins: int = 50,
                           **kw):
    """sigma_vs_time of the ensemble, computed on the device (no trajectory
    leaves the GPU): (t_centres, sigma_z, sigma_v, n_alive)."""
    s = ensemble_time_statistics(simulator, z0_arr, v0_arr, t_max, dt, time_bins, **kw)
    na = s["n_alive"]
    ok = na > 1
    sz = np.where(ok, np.sqrt(s["m2_z"] / np.maximum(na, 1)), np.nan)
    sv = np.where(ok, np.sqrt(s["m2_v"] / np.maximum(na, 1)), np.nan)
    return s["t_centres"], sz, sv, na


def temperature_vs_time_ensemble(simulator, z0_arr, v0_arr, t_max: float, dt: float = 1e-6, time_bins: int = 50,
                                 **kw):
    """temperature_vs_time of the ensemble on the device: (t_centres, T)."""
    from ..constants import k_B
    s = ensemble_time_statistics(simulator, z0_arr, v0_arr, t_max, dt, time_bins, **kw)
    na = s["n_alive"]
    v2 = (s["m2_v"] + na * s["mean_v"] ** 2) / np.maximum(na, 1)
    return s["t_centres"], np.where(na > 1, simulator.mol_data.mass * v2 / k_B, np.nan)


def density_vs_time_ensemble(simulator, z0_arr, v0_arr, t_max: float, dt: float = 1e-6, N_total: int = 2000,
                             time_bins: int = 50, **kw):
    """density_vs_time of the ensemble on the device: (t_centres, n0)."""
    tc, sz, _, na = sigma_vs_time_ensemble(simulator, z0_arr, v0_arr, t_max, dt, time_bins, **kw)
    return tc, _density(sz, na, len(np.ravel(z0_arr)), N_total)
