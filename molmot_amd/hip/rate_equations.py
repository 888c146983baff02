"""Batched 1D rate-equation force on MI355X.

API shape of molmot_cuda/rate_equations_cuda.py:448-765 (solve_rate_equations_batch,
force_vs_z/v, force_map_2d, spring_constant, damping_coefficient,
optimize_parameters), each available under the ``*_hip`` name and the
reference's ``*_cuda`` name.  Host numpy arrays in, host numpy arrays out;
``*_device`` variants take and return torch device tensors without host
round trips.
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import torch

from . import _lib
from .plan import HIPMolData


def _dev_f64(x, device):
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64)).to(device, non_blocking=False)


def rate_eq_map_device(plan: HIPMolData, v: torch.Tensor, z: torch.Tensor, B_gradient: float,
                       want_pop: bool = True, want_scatter: bool = True, out=None, stream=None):
    """F (nv*nz,), pop (nv*nz, n_g) | None, R (nv*nz,) | None on the device,
    index iv*nz + iz.  ``v``, ``z``: contiguous float64 cuda tensors."""
    if v.dtype != torch.float64 or z.dtype != torch.float64 or not v.is_cuda or not z.is_cuda:
        raise TypeError("v and z must be float64 cuda tensors")
    v = v.contiguous()
    z = z.contiguous()
    n = v.numel() * z.numel()
    if out is None:
        F = torch.empty(n, dtype=torch.float64, device=v.device)
        P = torch.empty((n, plan.n_ground), dtype=torch.float64, device=v.device) if want_pop else None
        R = torch.empty(n, dtype=torch.float64, device=v.device) if want_scatter else None
    else:
        F, P, R = out
    rc = _lib.lib().mm_rate_eq_map(plan.handle, _lib.ptr(v), v.numel(), _lib.ptr(z), z.numel(),
                                   float(B_gradient), _lib.ptr(F), _lib.ptr(P), _lib.ptr(R),
                                   _lib.stream_handle(stream))
    _lib.check(rc, "mm_rate_eq_map")
    return F, P, R


def rate_eq_points_device(plan: HIPMolData, v: torch.Tensor, z: torch.Tensor, B_gradient: float,
                          want_pop: bool = True, want_scatter: bool = True, stream=None):
    """Force at scattered points (v[i], z[i]) on the device."""
    if v.shape != z.shape:
        raise ValueError("v and z must have the same shape")
    v = v.contiguous().reshape(-1)
    z = z.contiguous().reshape(-1)
    n = v.numel()
    F = torch.empty(n, dtype=torch.float64, device=v.device)
    P = torch.empty((n, plan.n_ground), dtype=torch.float64, device=v.device) if want_pop else None
    R = torch.empty(n, dtype=torch.float64, device=v.device) if want_scatter else None
    rc = _lib.lib().mm_rate_eq_points(plan.handle, n, _lib.ptr(v), _lib.ptr(z), float(B_gradient),
                                      _lib.ptr(F), _lib.ptr(P), _lib.ptr(R), _lib.stream_handle(stream))
    _lib.check(rc, "mm_rate_eq_points")
    return F, P, R


def solve_rate_equations_batch(mol_data, beams, v_arr, z_arr, B_gradient: float,
                               Gamma_eff_factor: float = 1.0,
                               cuda_data: Optional[HIPMolData] = None,
                               hbar: Optional[float] = None) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """F (N_v, N_z) [N], pop (N_v, N_z, n_g), R (N_v, N_z) [1/s]
    (molmot_cuda/rate_equations_cuda.py:448-544)."""
    plan = cuda_data if cuda_data is not None else HIPMolData(mol_data, beams, Gamma_eff_factor, hbar=hbar)
    dev = torch.device("cuda", torch.cuda.current_device())
    v_arr = np.atleast_1d(np.asarray(v_arr, dtype=np.float64))
    z_arr = np.atleast_1d(np.asarray(z_arr, dtype=np.float64))
    F, P, R = rate_eq_map_device(plan, _dev_f64(v_arr, dev), _dev_f64(z_arr, dev), B_gradient)
    nv, nz = v_arr.size, z_arr.size
    return (F.cpu().numpy().reshape(nv, nz), P.cpu().numpy().reshape(nv, nz, plan.n_ground),
            R.cpu().numpy().reshape(nv, nz))


def force_vs_z_hip(mol_data, beams, z_arr, v: float, B_gradient: float, Gamma_eff_factor: float = 1.0,
                   cuda_data=None, hbar=None):
    """(molmot_cuda/rate_equations_cuda.py:551-594)"""
    F, P, R = solve_rate_equations_batch(mol_data, beams, [v], z_arr, B_gradient, Gamma_eff_factor,
                                         cuda_data, hbar)
    return F[0], P[0], R[0]


def force_vs_v_hip(mol_data, beams, v_arr, z: float, B_gradient: float, Gamma_eff_factor: float = 1.0,
                   cuda_data=None, hbar=None):
    """(molmot_cuda/rate_equations_cuda.py:597-637)"""
    F, _, _ = solve_rate_equations_batch(mol_data, beams, v_arr, [z], B_gradient, Gamma_eff_factor,
                                         cuda_data, hbar)
    return F[:, 0]


def force_map_2d_hip(mol_data, beams, v_arr, z_arr, B_gradient: float, Gamma_eff_factor: float = 1.0,
                     cuda_data=None, hbar=None):
    """(molmot_cuda/rate_equations_cuda.py:640-677)"""
    return solve_rate_equations_batch(mol_data, beams, v_arr, z_arr, B_gradient, Gamma_eff_factor,
                                      cuda_data, hbar)


def spring_constant_hip(mol_data, beams, B_gradient: float, z0: float = 0.0, dz: float = 0.3e-3,
                        Gamma_eff_factor: float = 1.0, cuda_data=None, hbar=None) -> float:
    """k = -(F(z0+dz) - F(z0-dz)) / (2 dz) at v = 0 (molmot_cuda/rate_equations_cuda.py:680-722)."""
    F, _, _ = solve_rate_equations_batch(mol_data, beams, [0.0], [z0 - dz, z0 + dz], B_gradient,
                                         Gamma_eff_factor, cuda_data, hbar)
    return float(-(F[0, 1] - F[0, 0]) / (2.0 * dz))


def damping_coefficient_hip(mol_data, beams, B_gradient: float, z0: float = 0.0, dv: float = 0.1,
                            Gamma_eff_factor: float = 1.0, cuda_data=None, hbar=None) -> float:
    """alpha = -(F(+dv) - F(-dv)) / (2 dv) at z0 (molmot_cuda/rate_equations_cuda.py:725-765)."""
    F, _, _ = solve_rate_equations_batch(mol_data, beams, [-dv, dv], [z0], B_gradient,
                                         Gamma_eff_factor, cuda_data, hbar)
    return float(-(F[1, 0] - F[0, 0]) / (2.0 * dv))


def optimize_parameters_hip(mol_data, delta_arr, split_arr, s0: float, B_gradient: float,
                            dz: float = 0.3e-3, dv: float = 0.1, Gamma_eff_factor: float = 1.0,
                            verbose: bool = False, hbar=None) -> dict:
    """(detuning, split) scan of k and beta = alpha/m
    (molmot_cuda/rate_equations_cuda.py:768-894; molmot/mot/force_scan.py:191-275)."""
    from ..obe.fields import make_dc_mot_beams
    delta_arr = np.asarray(delta_arr, dtype=float)
    split_arr = np.asarray(split_arr, dtype=float)
    mass = float(mol_data.mass)
    K = np.zeros((delta_arr.size, split_arr.size))
    beta = np.zeros_like(K)
    dev = torch.device("cuda", torch.cuda.current_device())
    v_t = _dev_f64([-dv, 0.0, dv], dev)
    z_t = _dev_f64([-dz, 0.0, dz], dev)
    results = []
    for i, det in enumerate(delta_arr):
        for j, sp in enumerate(split_arr):
            plan = HIPMolData(mol_data, make_dc_mot_beams(mol_data, det, sp, s0), Gamma_eff_factor, hbar=hbar)
            F, _, _ = rate_eq_map_device(plan, v_t, z_t, B_gradient, want_pop=False, want_scatter=False)
            results.append((i, j, F, plan))
    for i, j, F, plan in results:
        f = F.cpu().numpy().reshape(3, 3)
        K[i, j] = -(f[1, 2] - f[1, 0]) / (2.0 * dz)
        beta[i, j] = -(f[2, 1] - f[0, 1]) / (2.0 * dv) / mass
    merit = K.copy()
    merit[K <= 0] = 0
    merit[beta <= 0] = 0
    ib = np.unravel_index(np.argmax(merit), merit.shape)
    if verbose:
        print(f"best delta={delta_arr[ib[0]]:.3f} split={split_arr[ib[1]]:.3f}")
    return {"delta_scan": delta_arr, "split_scan": split_arr, "K_map": K, "beta_map": beta,
            "delta_best": float(delta_arr[ib[0]]), "split_best": float(split_arr[ib[1]]),
            "k_best": float(K[ib]), "beta_best": float(beta[ib])}


# reference names (molmot_cuda/__init__.py:72-81)
force_vs_z_cuda = force_vs_z_hip
force_vs_v_cuda = force_vs_v_hip
force_map_2d_cuda = force_map_2d_hip
spring_constant_cuda = spring_constant_hip
damping_coefficient_cuda = damping_coefficient_hip
optimize_parameters_cuda = optimize_parameters_hip
