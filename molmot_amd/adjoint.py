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
from .hip.plan import HIPMolData
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


def force_and_jacobian(problem: ForceProblem, theta, v: torch.Tensor, z: torch.Tensor, stream=None):
    """F (n,) and J (n, 9) = dF/d[det0..det3, split, s0, B', z, v] on the device."""
    plan = problem.plan(theta)
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
