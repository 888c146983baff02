"""Batched standing-wave Lindblad force on MI355X.

Physics follows molmot/obe/obe_subdoppler.py:19-286 (multi-rotating-frame
Hamiltonian, couplings Omega0 sqrt(s0) |d| / 2 e^{+-i kz}, spontaneous decay
through the jump operators sqrt(Gamma_eff br) |g><e|, force
2 hbar k Im(rho_ge c_eg)).  Where the reference solves for the steady state at
fixed standing-wave phase and ignores v, ``lindblad_rk4`` propagates rho in
time with RK4 while the phase moves as phi0 + k v t, so the force depends on
velocity; at v = 0 its long-time limit is the reference's steady state.

Beam pairs are the reference's dicts {'freq', 'q_fwd', 'q_bwd', 's0'}
(obe_subdoppler.py:289-337).  ``pol`` (n_pairs, 2, 3) complex replaces the
integer polarisations by polarisation vectors over q (the coupling then uses
the complex dipole sum_q eps_q d_q).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
import torch

from ..constants import hbar as HBAR
from ..hip import _lib

PRECISION = {"fp64": 0, "fp32": 1, "fp32_mfma": 2}


def _tables(mol, beam_pairs, z_macro, B_gradient, Omega_Gamma, Gamma_eff_factor, pol):
    """Hd (N,), A, B (n_e, n_g) complex, all in units of Gamma; gam (n_e,),
    rep (n_g, n_e) in units of Gamma."""
    G = mol.Gamma
    ng, ne = mol.n_ground, mol.n_excited
    B_gauss = B_gradient * z_macro * 1e4
    Es = mol.energies + B_gauss * mol.zeeman_z_diag * G / (2 * np.pi)
    Ee_mean = np.mean(Es[ng:])
    lf = [bp["freq"] for bp in beam_pairs]
    Hd = np.zeros(ng + ne)
    for ig in range(ng):
        om = Ee_mean - Es[ig]
        Hd[ig] = -2 * np.pi * min((f - om for f in lf), key=abs)
    for ie in range(ne):
        Hd[ng + ie] = 2 * np.pi * (Es[ng + ie] - Ee_mean)
    A = np.zeros((ne, ng), dtype=complex)
    Bm = np.zeros((ne, ng), dtype=complex)
    d = mol.tdm
    for p, bp in enumerate(beam_pairs):
        Om = Omega_Gamma * G * np.sqrt(bp["s0"])
        for ig in range(ng):
            for ie in range(ne):
                for side, M in ((0, A), (1, Bm)):
                    if pol is None:
                        q = bp["q_fwd"] if side == 0 else bp["q_bwd"]
                        dv = abs(d[ig, ng + ie, q])
                    else:
                        dv = np.sum(np.asarray(pol)[p, side] * d[ig, ng + ie, :])
                    if abs(dv) <= 1e-10:
                        continue
                    M[ie, ig] += Om * dv / 2
    d2 = mol.d_squared
    rep = np.zeros((ng, ne))
    for ie in range(ne):
        total = sum(d2[ig, ie, q] for ig in range(ng) for q in range(3))
        if total < 1e-30:
            continue
        for ig in range(ng):
            for q in range(3):
                br = d2[ig, ie, q] / total
                if br >= 1e-15:
                    rep[ig, ie] += Gamma_eff_factor * br
    return Hd / G, A / G, Bm / G, rep.sum(0), rep


class LindbladClass:
    """Device tables of one system class (molecule, beams, field, Rabi
    frequency); instances differ by velocity and standing-wave phase."""

    def __init__(self, mol, beam_pairs, z_macro=0.0, B_gradient=0.0, Omega_Gamma=1.0,
                 Gamma_eff_factor=1.0, pol=None):
        _lib.require_gpu()
        self.mol = mol
        self.n_ground, self.n_excited = int(mol.n_ground), int(mol.n_excited)
        self.N = self.n_ground + self.n_excited
        Hd, A, B, gam, rep = _tables(mol, beam_pairs, z_macro, B_gradient, Omega_Gamma, Gamma_eff_factor, pol)
        self.Hd, self.A, self.B, self.gam, self.rep = Hd, A, B, gam, rep
        self.kfac = mol.k / mol.Gamma
        self.force_unit = 2.0 * HBAR * mol.k * mol.Gamma
        self.rate_unit = mol.Gamma * Gamma_eff_factor
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in
                (Hd, A.real, A.imag, B.real, B.imag, gam, rep)]
        dp = [a.ctypes.data_as(_lib._dp) for a in arrs]
        self._h = ctypes.c_void_p()
        rc = _lib.lib().mm_lindblad_class_create(ctypes.byref(self._h), self.n_ground, self.n_excited, *dp,
                                                 self.force_unit, self.rate_unit)
        _lib.check(rc, "mm_lindblad_class_create")

    def set_lanes(self, precision: str, lanes: int) -> None:
        """Lanes per instance of the VALU propagator for precision "fp64" or
        "fp32": 64, 128, 256, or 0 for the default."""
        _lib.check(_lib.lib().mm_lindblad_class_set_lanes(self._h, PRECISION[precision], int(lanes)),
                   "mm_lindblad_class_set_lanes")

    def lanes(self, precision: str) -> int:
        n = _lib.lib().mm_lindblad_class_get_lanes(self._h, PRECISION[precision])
        if n < 0:
            _lib.check(n, "mm_lindblad_class_get_lanes")
        return n

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.check(_lib.lib().mm_lindblad_class_destroy(self._h), "mm_lindblad_class_destroy")
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lindblad_rk4_device(cls: LindbladClass, kv: torch.Tensor, phi0: torch.Tensor,
                        dt: float, n_steps: int, avg_start: int, precision: str = "fp64",
                        want_rho: bool = False, stream=None):
    """Per-instance (F [N], R [1/s], rho_final | None) on the device; kv = k v / Gamma."""
    kv = kv.contiguous().to(torch.float64)
    phi0 = phi0.contiguous().to(torch.float64)
    n = kv.numel()
    F = torch.empty(n, dtype=torch.float64, device=kv.device)
    R = torch.empty(n, dtype=torch.float64, device=kv.device)
    rdt = torch.float64 if precision == "fp64" else torch.float32
    rho = torch.empty((n, cls.N, cls.N, 2), dtype=rdt, device=kv.device) if want_rho else None
    rc = _lib.lib().mm_lindblad_rk4(cls._h, n, _lib.ptr(kv), _lib.ptr(phi0), float(dt), int(n_steps),
                                    int(avg_start), PRECISION[precision], _lib.ptr(F), _lib.ptr(R),
                                    _lib.ptr(rho), _lib.stream_handle(stream))
    _lib.check(rc, "mm_lindblad_rk4")
    if rho is not None:
        rho = torch.complex(rho[..., 0].to(torch.float64), rho[..., 1].to(torch.float64))
    return F, R, rho


def lindblad_rk4_uneven_device(cls16: LindbladClass, kv16: torch.Tensor, phi16: torch.Tensor,
                               cls40: LindbladClass, kv40: torch.Tensor, phi40: torch.Tensor, dt: float,
                               n_steps: int, avg_start: int, precision: str = "fp32_mfma", want_rho: bool = False,
                               stream=None):
    """A 12+4 and a 36+4 batch propagated in ONE launch (uneven instance
    sizes): ((F16, R16, rho16 | None), (F40, R40, rho40 | None)), identical to
    two lindblad_rk4_device calls.  precision "fp64" (VALU for both) or
    "fp32_mfma" (16-level part on MFMA, 40-level part on FP32 VALU)."""
    if precision not in ("fp64", "fp32_mfma"):
        raise ValueError("uneven launch precision is 'fp64' or 'fp32_mfma'")
    outs = []
    for cls, kv, ph in ((cls16, kv16, phi16), (cls40, kv40, phi40)):
        kv = kv.contiguous().to(torch.float64)
        ph = ph.contiguous().to(torch.float64)
        n = kv.numel()
        rdt = torch.float64 if precision == "fp64" else torch.float32
        outs.append((kv, ph, torch.empty(n, dtype=torch.float64, device=kv.device),
                     torch.empty(n, dtype=torch.float64, device=kv.device),
                     torch.empty((n, cls.N, cls.N, 2), dtype=rdt, device=kv.device) if want_rho else None))
    (k1, p1, F1, R1, r1), (k2, p2, F2, R2, r2) = outs
    rc = _lib.lib().mm_lindblad_rk4_uneven(cls16._h, k1.numel(), _lib.ptr(k1), _lib.ptr(p1), _lib.ptr(F1),
                                           _lib.ptr(R1), _lib.ptr(r1), cls40._h, k2.numel(), _lib.ptr(k2),
                                           _lib.ptr(p2), _lib.ptr(F2), _lib.ptr(R2), _lib.ptr(r2), float(dt),
                                           int(n_steps), int(avg_start), PRECISION[precision],
                                           _lib.stream_handle(stream))
    _lib.check(rc, "mm_lindblad_rk4_uneven")
    res = []
    for F, R, rho in ((F1, R1, r1), (F2, R2, r2)):
        if rho is not None:
            rho = torch.complex(rho[..., 0].to(torch.float64), rho[..., 1].to(torch.float64))
        res.append((F, R, rho))
    return tuple(res)


def obe_force_rk4(mol, beam_pairs, v_arr, z_macro: float = 0.0, B_gradient: float = 0.0,
                  Omega_Gamma: float = 1.0, n_spatial: int = 64, t_total: float = 200 * 2 * np.pi,
                  dt: float = 0.05, avg_from: float = 0.5, Gamma_eff_factor: float = 1.0, pol=None,
                  precision: str = "fp64"):
    """Velocity-dependent standing-wave force: for each v, the RK4 force
    averaged over n_spatial phases and over the second part of the run.
    Times in units of 1/Gamma.  Returns (F (n_v,) [N], R (n_v,) [1/s])."""
    cls = LindbladClass(mol, beam_pairs, z_macro, B_gradient, Omega_Gamma, Gamma_eff_factor, pol)
    v = np.atleast_1d(np.asarray(v_arr, dtype=np.float64))
    phis = np.linspace(0.0, 2 * np.pi, n_spatial, endpoint=False)
    V, P = np.meshgrid(v, phis, indexing="ij")
    dev = torch.device("cuda", torch.cuda.current_device())
    n_steps = int(round(t_total / dt))
    avg_start = int(avg_from * n_steps)
    F, R, _ = lindblad_rk4_device(cls, torch.from_numpy((V * cls.kfac).ravel()).to(dev),
                                  torch.from_numpy(P.ravel()).to(dev), dt, n_steps, avg_start, precision)
    F = F.cpu().numpy().reshape(v.size, n_spatial).mean(1)
    R = R.cpu().numpy().reshape(v.size, n_spatial).mean(1)
    return F, R


def liouvillian_device(cls: LindbladClass, phi0: torch.Tensor, trace_row: bool = True, stream=None):
    """(n, N^2, N^2) complex128 Liouvillians (units of Gamma, column-major vec)
    at the standing-wave phases phi0; with trace_row the last row is the
    trace constraint."""
    phi0 = phi0.contiguous().to(torch.float64)
    n = phi0.numel()
    N2 = cls.N * cls.N
    M = torch.empty((n, N2, N2, 2), dtype=torch.float64, device=phi0.device)
    rc = _lib.lib().mm_lindblad_liouvillian(cls._h, n, _lib.ptr(phi0), int(bool(trace_row)), _lib.ptr(M),
                                            _lib.stream_handle(stream))
    _lib.check(rc, "mm_lindblad_liouvillian")
    return torch.view_as_complex(M)


# Items whose LU has min|U_ii| / max|U_ii| below this are treated as
# rank-deficient (several stationary states, e.g. at standing-wave nodes where
# every coupling cancels) and solved for the minimum-norm stationary state.
RANK_DEFICIENT_PIVOT_RATIO = 1e-10


def steady_state_device(cls: LindbladClass, phi0: torch.Tensor, clip_negative: bool = True, stream=None):
    """Steady-state density matrices at the given standing-wave phases.

    The kernel builds the Liouvillians with the trace constraint in the last
    row (molmot_cuda/lindblad_cuda.py:199-258) and the library's batched
    complex LU kernel (mm_zgesv_batched) solves them.  Items whose pivots show a rank-deficient system
    get the minimum-norm least-squares solution of the trace-augmented system
    with numpy's default cut-off (eps * (N^2 + 1) * sigma_max), which is what
    the reference's CPU path computes for every item
    (molmot/obe/obe_subdoppler.py:144-165).  Then Hermitisation, optional
    clipping of negative eigenvalues (obe_subdoppler.py:166-183) and trace
    normalisation.  -> (n, N, N) complex128."""
    phi0 = phi0.contiguous().to(torch.float64)
    n = phi0.numel()
    N = cls.N
    N2 = N * N
    M = liouvillian_device(cls, phi0, True, stream)
    rhs = torch.zeros((n, N2, 1), dtype=torch.complex128, device=phi0.device)
    rhs[:, -1, 0] = 1.0
    # LU with partial pivoting in the library's batched kernel, one
    # workgroup per Liouvillian (M is overwritten by its factors)
    info = torch.zeros(n, dtype=torch.int32, device=phi0.device)
    umin = torch.empty(n, dtype=torch.float64, device=phi0.device)
    umax = torch.empty(n, dtype=torch.float64, device=phi0.device)
    rc = _lib.lib().mm_zgesv_batched(N2, 1, n, _lib.ptr(M), _lib.ptr(rhs), _lib.ptr(info), _lib.ptr(umin),
                                     _lib.ptr(umax), _lib.stream_handle(stream))
    _lib.check(rc, "mm_zgesv_batched")
    x = rhs[..., 0]
    del M
    bad = (umin < RANK_DEFICIENT_PIVOT_RATIO * umax) | (info != 0) | ~torch.isfinite(x).all(-1)
    idx = torch.nonzero(bad).flatten()
    if idx.numel():
        L = liouvillian_device(cls, phi0[idx], False, stream)
        tr = torch.zeros((idx.numel(), 1, N2), dtype=torch.complex128, device=phi0.device)
        tr[:, 0, torch.arange(N, device=phi0.device) * (N + 1)] = 1.0
        A = torch.cat([L, tr], 1)
        b = torch.zeros((idx.numel(), N2 + 1, 1), dtype=torch.complex128, device=phi0.device)
        b[:, -1, 0] = 1.0
        eps = float(np.finfo(np.float64).eps)
        x[idx] = (torch.linalg.pinv(A, rtol=eps * (N2 + 1)) @ b)[..., 0]
    rho = x.reshape(n, N, N).transpose(1, 2)          # column-major vec -> rho[i, j]
    rho = 0.5 * (rho + rho.conj().transpose(1, 2))
    if clip_negative:
        w, U = torch.linalg.eigh(rho)
        rho = (U * w.clamp(min=0).to(U.dtype).unsqueeze(1)) @ U.conj().transpose(1, 2)
    tr = torch.diagonal(rho, dim1=1, dim2=2).sum(-1)
    tr = torch.where(tr.abs() > 1e-15, tr, torch.ones_like(tr))
    return rho / tr[:, None, None]


def _force_rate(cls: LindbladClass, rho: torch.Tensor, phi: torch.Tensor):
    """Per-instance force (N) and excited population from rho (obe_subdoppler.py:186-228)."""
    ng = cls.n_ground
    A = torch.as_tensor(cls.A, device=rho.device)
    B = torch.as_tensor(cls.B, device=rho.device)
    e = torch.exp(1j * phi.to(torch.float64)).to(torch.complex128)
    W = e[:, None, None] * A[None] - (1.0 / e)[:, None, None] * B[None]        # (n, n_e, n_g)
    blk = rho[:, :ng, ng:]                                                      # rho[g, e]
    F = cls.force_unit * torch.imag((blk * W.transpose(1, 2)).sum((1, 2)))
    pe = torch.real(torch.diagonal(rho, dim1=1, dim2=2)[:, ng:]).sum(-1)
    return F, pe


def obe_force_subdoppler(mol, beam_pairs, v, z_macro, B_gradient, Omega_Gamma, n_spatial: int = 16,
                         Gamma_eff_factor: float = 1.0, pol=None, clip_negative: bool = True):
    """(F [N], rho_avg (N, N), R [1/s]) with the reference signature
    (molmot/obe/obe_subdoppler.py:231-286).  As in the reference, the
    steady state does not depend on v."""
    F, rho, R = obe_force_subdoppler_hip(mol, beam_pairs, [v], z_macro, B_gradient, Omega_Gamma, n_spatial,
                                         Gamma_eff_factor, pol, clip_negative)
    return float(F[0]), rho[0], float(R[0])


def obe_force_subdoppler_hip(mol, beam_pairs, v_arr, z_macro=0.0, B_gradient=0.0, Omega_Gamma=1.0,
                             n_spatial: int = 16, Gamma_eff_factor: float = 1.0, pol=None,
                             clip_negative: bool = False):
    """(F (n_v,), rho_avg (n_v, N, N), R (n_v,)) with the GPU reference's
    signature (molmot_cuda/lindblad_cuda.py:448-538): n_spatial steady
    states, averaged; every velocity gets the same (v-independent) result."""
    cls = LindbladClass(mol, beam_pairs, z_macro, B_gradient, Omega_Gamma, Gamma_eff_factor, pol)
    dev = torch.device("cuda", torch.cuda.current_device())
    phis = torch.as_tensor(np.linspace(0, 2 * np.pi, n_spatial, endpoint=False), device=dev)
    rho = steady_state_device(cls, phis, clip_negative)
    F, pe = _force_rate(cls, rho, phis)
    rho_avg = rho.mean(0).cpu().numpy()
    Fm = float(F.mean().item())
    R = cls.rate_unit * float(np.sum(np.real(np.diag(rho_avg))[cls.n_ground:]))
    nv = len(np.atleast_1d(v_arr))
    return np.full(nv, Fm), np.repeat(rho_avg[None], nv, 0), np.full(nv, max(R, 0.0))


obe_force_subdoppler_cuda = obe_force_subdoppler_hip
