"""Ensemble estimators over quantum-jump trajectories
(molmot/obe/diffusion.py:35-239): momentum diffusion from the growth of the
velocity variance, damping rate from the velocity autocorrelation, photon
scattering rate.  ``results`` is a list of SSEResult (SI units); the *_device
variants take the (n, n_save, 3) velocity tensor of mcwf_ensemble_device
and reduce on the GPU."""
from __future__ import annotations

import numpy as np

from ..constants import k_B


def _vstats(results, axis):
    times = results[0].times
    nt = len(times)
    s1 = np.zeros(nt)
    s2 = np.zeros(nt)
    for r in results:
        n = min(nt, len(r.velocities))
        v = r.velocities[:n, axis]
        s1[:n] += v
        s2[:n] += v * v
    n = len(results)
    return times, s1 / n, s2 / n


def _diffusion_fit(times, var, mass, t_min):
    i0 = np.searchsorted(times, t_min) if t_min is not None else max(1, len(times) // 4)
    if i0 >= len(times) - 1:
        return 0.0, 0.0, float("nan")
    tf, vf = times[i0:], var[i0:]
    if len(tf) < 2:
        return 0.0, 0.0, float("nan")
    c = np.polyfit(tf, vf, 1)
    res = vf - np.polyval(c, tf)
    return mass ** 2 * c[0] / 2.0, mass ** 2 * np.std(res) / (2.0 * np.sqrt(len(tf))), float("nan")


def compute_diffusion_from_ensemble(results, mass, axis=2, t_min=None):
    """(D, D_err, nan): D = m^2 d Var(v)/dt / 2 from a linear fit after t_min
    (default: skip the first quarter) (diffusion.py:35-125)."""
    if len(results) == 0:
        return 0.0, 0.0, float("nan")
    t, m1, m2 = _vstats(results, axis)
    return _diffusion_fit(t, m2 - m1 ** 2, mass, t_min)


def compute_diffusion_from_ensemble_device(v, times, mass, axis=2, t_min=None):
    """Same estimator on a device tensor v (n, n_save, 3) [m/s] with times (n_save,) [s]."""
    x = v[..., axis]
    var = (x * x).mean(0) - x.mean(0) ** 2
    return _diffusion_fit(np.asarray(times), var.cpu().numpy(), mass, t_min)


def compute_diffusion_temperature(D, mass, beta):
    """T = D / (m beta k_B) (diffusion.py:128-151)."""
    if abs(beta) < 1e-30:
        return float("inf")
    return D / (mass * beta * k_B)


def _damping_fit(times, v_auto, v0_sq):
    if v0_sq < 1e-20:
        return 0.0
    a = v_auto / v0_sq
    ok = a > 0.1
    if np.sum(ok) < 3:
        return 0.0
    c = np.polyfit(times[ok], np.log(a[ok]), 1)
    return max(-c[0], 0.0)


def estimate_damping_rate(results, mass, axis=2, v_range=None):
    """beta from <v(t) v(0)> / <v(0)^2> ~ exp(-beta t) (diffusion.py:154-211)."""
    if len(results) == 0:
        return 0.0
    times = results[0].times
    nt = len(times)
    auto = np.zeros(nt)
    v0s = 0.0
    for r in results:
        n = min(nt, len(r.velocities))
        v = r.velocities[:n, axis]
        v0s += v[0] ** 2
        auto[:n] += v[0] * v[:n]
    n = len(results)
    return _damping_fit(times, auto / n, v0s / n)


def estimate_damping_rate_device(v, times, axis=2):
    x = v[..., axis]
    auto = (x[:, :1] * x).mean(0).cpu().numpy()
    return _damping_fit(np.asarray(times), auto, float((x[:, 0] ** 2).mean()))


def compute_scattering_rate(results, t_min=None):
    """(mean, standard error) of the per-trajectory photon count rate
    (diffusion.py:214-239).  Trajectories run without a jump log (empty
    jump_times, as the GPU ensemble returns them) count photons_scattered
    over the whole run when t_min is None; with t_min they raise ValueError
    (the photons after t_min are unknown)."""
    rates = []
    for r in results:
        if t_min is not None:
            if not r.jump_times and r.photons_scattered:
                raise ValueError("compute_scattering_rate: t_min needs the jump times; run the ensemble "
                                 "with log_jumps=True (only the photon count was kept)")
            n = sum(1 for t in r.jump_times if t > t_min)
            dt = r.times[-1] - t_min
        else:
            n = len(r.jump_times) if (r.jump_times or not r.photons_scattered) else r.photons_scattered
            dt = r.times[-1] - r.times[0]
        if dt > 0:
            rates.append(n / dt)
    if not rates:
        return 0.0, 0.0
    rates = np.array(rates)
    return float(np.mean(rates)), float(np.std(rates) / np.sqrt(len(rates)))
