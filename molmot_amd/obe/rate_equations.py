"""Single-point rate-equation solve with the reference signature
(molmot/obe/rate_equations.py:23-191, dispatch :205-218), evaluated by the
gfx950 kernel.  Batched callers should use molmot_amd.hip directly."""
from __future__ import annotations

import numpy as np
import torch

from ..hip.plan import HIPMolData
from ..hip.rate_equations import rate_eq_points_device

_PLAN_CACHE: dict = {}


def _plan(mol_data, beams, Gamma_eff_factor, hbar):
    key = (id(mol_data), tuple((b.kdir_z, b.freq_offset, int(b.polarization), b.s0) for b in beams),
           float(Gamma_eff_factor), hbar)
    p = _PLAN_CACHE.get(key)
    if p is None:
        if len(_PLAN_CACHE) > 64:
            _PLAN_CACHE.clear()
        p = HIPMolData(mol_data, beams, Gamma_eff_factor, hbar=hbar)
        _PLAN_CACHE[key] = p
    return p


def solve_rate_equations(mol_data, beams, v: float, z: float, B_gradient: float,
                         Gamma_eff_factor: float = 1.0, hbar: float | None = None):
    """(force [N], populations (n_ground,), R_scatter [1/s]) at one (v, z)."""
    plan = _plan(mol_data, beams, Gamma_eff_factor, hbar)
    dev = torch.device("cuda", torch.cuda.current_device())
    vt = torch.tensor([float(v)], dtype=torch.float64, device=dev)
    zt = torch.tensor([float(z)], dtype=torch.float64, device=dev)
    F, P, R = rate_eq_points_device(plan, vt, zt, B_gradient)
    return float(F.item()), P[0].cpu().numpy(), float(R.item())


solve_rate_equations_auto = solve_rate_equations
solve_rate_equations_jit = solve_rate_equations
