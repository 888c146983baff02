"""Gradient-driven laser/field parameter search.

The reference tunes a MOT by scanning a (detuning, split) grid and keeping the
best spring constant with positive damping (optimize_parameters,
molmot/mot/force_scan.py:191-301; molmot_cuda/rate_equations_cuda.py:768-894)
or by a black-box Bayesian search.  Here the objective's exact gradient comes
from the adjoint kernels (molmot_amd.adjoint), so the search follows it:
projected Adam ascent in scaled coordinates over any linear subspace of the
7 parameters theta = [det0..det3 (Hz), split (Hz), s0 scale, B' (T/m)],
restarted from a few points and keeping the best.
"""
from __future__ import annotations

import math
from typing import Callable, Dict, Optional, Sequence

import numpy as np
import torch

from .adjoint import N_THETA, ForceProblem, damping_coefficient_t, spring_constant_t
from .obe.fields import make_dc_mot_beams


def gradient_search(theta0, objective: Callable, directions, scales,
                    lower=None, upper=None, steps: int = 60, lr: float = 0.05) -> Dict:
    """Maximise objective(theta) over theta = theta0 + (u * scales) @ directions.

    directions (k, 7), scales (k,): one row per free variable; u starts at 0
    and is kept within [lower, upper] (in units of scales) by projection after
    each Adam step.  ``objective`` returns either a differentiable scalar
    (ascended and ranked as is) or a triple (ascent, merit, kind): the step
    follows the gradient of ``ascent`` and the best point is the one of
    largest float ``merit`` (first visited on ties); ``kind`` names the
    quantity being ascended, each kind scaled by the magnitude of its first
    value.  Returns {"theta", "value", "merit", "u", "history": [(merit, u)]}
    for the best point visited; "value" is its ascent objective."""
    th0 = torch.as_tensor(np.asarray(theta0, dtype=np.float64))
    D = torch.as_tensor(np.asarray(directions, dtype=np.float64))
    sc = torch.as_tensor(np.asarray(scales, dtype=np.float64))
    k = D.shape[0]
    lo = None if lower is None else torch.as_tensor(np.asarray(lower, dtype=np.float64))
    hi = None if upper is None else torch.as_tensor(np.asarray(upper, dtype=np.float64))
    u = torch.zeros(k, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([u], lr=lr, eps=1e-12)
    best_m, best_v, best_u, hist = -math.inf, -math.inf, u.detach().clone(), []
    norm = {}   # objective scale per kind: physical units (N/m ~ 1e-20) would vanish under Adam's eps
    for _ in range(steps):
        opt.zero_grad()
        theta = th0 + (u * sc) @ D
        out = objective(theta)
        f, merit, kind = (out if isinstance(out, tuple) else (out, None, None))
        fv = float(f.detach())
        mv = fv if merit is None else float(merit)
        hist.append((mv, u.detach().clone().numpy()))
        if mv > best_m:
            best_m, best_v, best_u = mv, fv, u.detach().clone()
        if kind not in norm:
            norm[kind] = 1.0 / max(abs(fv), 1e-300)
        (-f * norm[kind]).backward()
        opt.step()
        with torch.no_grad():
            if lo is not None:
                u.copy_(torch.maximum(u, lo))
            if hi is not None:
                u.copy_(torch.minimum(u, hi))
    theta = (th0 + (best_u * sc) @ D).numpy()
    return {"theta": theta, "value": best_v, "merit": best_m, "u": best_u.numpy(), "history": hist}


def constrained_k(theta, prob: ForceProblem, mass: float, dz: float, dv: float):
    """(ascent, merit, kind) of the reference's selection rule (force_scan.py:256-262):
    merit = k where k > 0 and beta = alpha / m > 0, else 0.  The ascent
    follows k where the damping is positive and the damping itself where it
    is not, so a search leaves an anti-damping region before it climbs k."""
    k = spring_constant_t(theta, prob, dz=dz)
    alpha = damping_coefficient_t(theta, prob, dv=dv)
    kv, av = float(k.detach()), float(alpha.detach())
    merit = kv if (kv > 0.0 and av / mass > 0.0) else 0.0
    return (k, merit, "k") if av > 0.0 else (alpha, merit, "alpha")


def optimize_parameters_grad(mol_data, B_gradient: float, delta_range=(-2.0, 0.0), split_range=(0.0, 2.0),
                             s0: float = 1.0, starts: Optional[Sequence] = None, steps: int = 60,
                             lr: float = 0.05, independent: bool = False, dz: float = 0.3e-3,
                             dv: float = 0.1) -> Dict:
    """DC MOT: maximise the spring constant k over the detuning and split (in
    units of Gamma, within the given ranges) by gradient ascent, as
    optimize_parameters does by a grid scan (molmot/mot/force_scan.py:191-301).

    The reference's rule decides the optimum: only points with k > 0 and
    damping beta = alpha/m > 0 count (force_scan.py:256-262); the ascent
    follows k where alpha > 0 and alpha where it is not (constrained_k).
    independent=True frees the four frequency components' detunings
    separately (five variables with the split), a search no grid of the
    reference's size reaches.  Returns
    {"delta_best", "split_best", "theta", "k_best", "beta_best",
    "evaluations", "feasible"}; k and beta = alpha/m at the optimum as in
    the reference.  If no visited point has k > 0 and beta > 0 the first
    start is returned with feasible=False, as the reference's all-zero merit
    map returns its first grid point."""
    G = float(mol_data.Gamma)
    hz = G / (2.0 * math.pi)
    d_mid = 0.5 * (delta_range[0] + delta_range[1])
    s_mid = 0.5 * (split_range[0] + split_range[1])
    prob = ForceProblem(mol_data, make_dc_mot_beams(mol_data, d_mid, s_mid, s0))
    theta_c = prob.theta0(B_gradient).numpy()
    if independent:
        D = np.zeros((5, N_THETA))
        for c in range(4):
            D[c, c] = 1.0
        D[4, 4] = 1.0
        half = [0.5 * (delta_range[1] - delta_range[0])] * 4 + [0.5 * (split_range[1] - split_range[0])]
    else:
        D = np.zeros((2, N_THETA))
        D[0, 0:4] = 1.0          # common detuning
        D[1, 4] = 1.0            # split
        half = [0.5 * (delta_range[1] - delta_range[0]), 0.5 * (split_range[1] - split_range[0])]
    half = np.asarray(half)
    scales = half * hz          # u in [-1, 1] spans the range
    if starts is None:
        starts = [(a, b) for a in (-0.5, 0.0, 0.5) for b in (-0.5, 0.0, 0.5)]
    mass = float(mol_data.mass)
    objective = lambda th: constrained_k(th, prob, mass, dz, dv)  # noqa: E731
    best = None
    first = None
    evals = 0
    for st in starts:
        st = np.asarray(st, dtype=float)
        u0 = np.zeros(D.shape[0])
        if independent:
            u0[:4] = st[0]
            u0[4] = st[1]
        else:
            u0[:] = st
        th_start = theta_c + (u0 * scales) @ D
        r = gradient_search(th_start, objective, D, scales, lower=-1.0 - u0, upper=1.0 - u0, steps=steps, lr=lr)
        evals += steps
        if first is None:
            first = th_start
        if best is None or r["merit"] > best["merit"]:
            best = r
    feasible = best["merit"] > 0.0
    th = best["theta"] if feasible else first
    k = float(spring_constant_t(torch.tensor(th), prob, dz=dz))
    alpha = float(damping_coefficient_t(torch.tensor(th), prob, dv=dv))
    return {"delta_best": d_mid + float(np.mean(th[0:4] - theta_c[0:4])) / hz,
            "split_best": s_mid + float(th[4] - theta_c[4]) / hz,
            "theta": th, "k_best": k, "beta_best": alpha / mass, "evaluations": evals, "feasible": feasible}
