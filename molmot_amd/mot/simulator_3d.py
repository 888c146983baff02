"""3D MOT with the full ground-state Zeeman Hamiltonian, on MI355X.

Drop-in for molmot/mot/simulator_3d.py (Beam3D, make_6beam_config,
rate_eq_force_3d, MOTSimulator3D, compute_3d_force_map) and
molmot_cuda/simulator_3d_cuda.py (rate_eq_force_3d_batch,
compute_3d_force_map_cuda, simulate_trajectory_3d_cuda,
simulate_ensemble_3d_cuda).  Every force evaluation runs in the gfx950
kernels of molmot_amd/csrc/sim3d.hip; trajectories run as one persistent
leapfrog launch instead of one launch per step.

Degenerate Zeeman spectra: the reference pairs eigenvalue i (ascending) with
bare energy i and weights transitions by |U^+ |d||^2, so wherever eigenvalues
coincide its force depends on which orthonormal basis of the degenerate
eigenspace the eigensolver happens to return (for the SrOH tables the ground
block has just two eigenvalues +-g|B|/2, each six-fold).  Two bases are
available:

* ``basis="spin"`` (the default wherever it applies): when the ground Zeeman
  blocks are s times the spin operators of a sum of spin-1/2 representations
  ([Zx, Zy] = i s Zz cyclically and Zx^2 + Zy^2 + Zz^2 = 3 s^2 / 4, to a
  relative 1e-9; the SrOH tables hold to 2e-12), the eigenvectors of
  H_Z = B.Z are R(B) W: W the eigenvectors of Zz (numpy's eigh, once), R the
  spin rotation taking z to B / |B| about z x B.  This is a closed form (no
  per-point diagonalisation: one lane per point) and a basis that varies
  continuously with the field direction except at B = -z; on the z axis
  (B along +z) it is the eigh basis of the reference, and at B = 0 the
  identity, as eigh returns for the zero matrix.
* ``basis="jacobi"``: per-point complex Jacobi in LDS, 16 lanes per point,
  for any Hermitian tables; the basis its sweeps converge to.

For non-degenerate spectra both equal the reference's result.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np
import torch

from ..constants import hbar
from ..hip import _lib


class Beam3D:
    """One beam: unit direction k_hat, absolute frequency [Hz], lab-frame
    polarisation q_lab (0 sigma-, 1 pi, 2 sigma+), s0, 1/e^2 radius [m] or
    None (simulator_3d.py:31-59)."""

    def __init__(self, k_hat, freq, q_lab, s0, beam_radius=None):
        k = np.asarray(k_hat, dtype=float)
        self.k_hat = k / np.linalg.norm(k)
        self.freq = freq
        self.q_lab = q_lab
        self.s0 = s0
        self.beam_radius = beam_radius


_DIRS = [np.array([1, 0, 0]), np.array([-1, 0, 0]), np.array([0, 1, 0]), np.array([0, -1, 0]),
         np.array([0, 0, 1]), np.array([0, 0, -1])]


def make_6beam_config(mol_data, beam_pairs, beam_radius=10e-3):
    """Per beam pair, beams along +x -x +y -y +z -z; forward (+) beams carry
    q_fwd, backward beams q_bwd (simulator_3d.py:62-108)."""
    beams = []
    for bp in beam_pairs:
        for i, d in enumerate(_DIRS):
            beams.append(Beam3D(d, bp["freq"], bp["q_fwd"] if i % 2 == 0 else bp["q_bwd"], bp["s0"], beam_radius))
    return beams


def _decompose_polarization_local(q_lab, B_hat):
    """Squared Wigner d^1(theta) column for q_lab with theta the polar angle
    of B_hat, normalised (simulator_3d.py:111-165); host evaluation."""
    n = np.linalg.norm(B_hat)
    if n < 1e-15:
        return np.full(3, 1.0 / 3.0)
    ct = np.clip(B_hat[2] / n, -1.0, 1.0)
    st = np.sin(np.arccos(ct))
    a, b, c = (1 + ct) / 2, (1 - ct) / 2, st / np.sqrt(2)
    m = np.array([[a * a, c * c, b * b], [c * c, ct * ct, c * c], [b * b, c * c, a * a]])
    w = m[:, q_lab].copy()
    s = w.sum()
    return w / s if s > 1e-15 else w


def spin_basis_tables(mol_data, tol=1e-9):
    """(lam0 (n_g,), sd (n_g, 3 n_e, 6), inv_tot (n_e,)) of the spin-rotation
    basis, or None when the ground Zeeman blocks are not s times spin-1/2
    operators to relative ``tol``.  lam0, W = eigh(Zz) (Hermitised ground
    blocks); sd[g, c] = (W^H|d|, W^H (2/s) Zx|d|, W^H (2/s) Zy|d|)[g, c] as
    (re, im) pairs, |d| the n_g x 3 n_e table |tdm[g, n_g + e, q]| at column
    3 e + q; inv_tot[e] = 1 / sum_{g,q} |d|^2 (0 where <= 1e-30)."""
    m = mol_data
    ng, ne = m.n_ground, m.n_excited
    Z = []
    for a in "xyz":
        z = np.asarray(getattr(m, f"zeeman_{a}"))[:ng, :ng].astype(complex)
        Z.append(0.5 * (z + z.conj().T))
    Zx, Zy, Zz = Z
    C = Zx @ Zx + Zy @ Zy + Zz @ Zz
    s2 = float(np.real(np.trace(C))) / (0.75 * ng)
    if not s2 > 0.0:
        return None
    sc = np.sqrt(s2)
    bad = max(np.max(np.abs(C - 0.75 * s2 * np.eye(ng))),
              np.max(np.abs(Zx @ Zy - Zy @ Zx - 1j * sc * Zz)),
              np.max(np.abs(Zy @ Zz - Zz @ Zy - 1j * sc * Zx)),
              np.max(np.abs(Zz @ Zx - Zx @ Zz - 1j * sc * Zy)))
    if not bad <= tol * s2:
        return None
    lam0, W = np.linalg.eigh(Zz)
    d = np.abs(np.asarray(m.tdm)[:ng, ng:ng + ne, :]).reshape(ng, 3 * ne).astype(complex)
    Wh = W.conj().T
    D0, Dx, Dy = Wh @ d, Wh @ ((2.0 / sc) * Zx) @ d, Wh @ ((2.0 / sc) * Zy) @ d
    sd = np.stack([D0.real, D0.imag, Dx.real, Dx.imag, Dy.real, Dy.imag], axis=-1)
    tot = (np.abs(d) ** 2).reshape(ng, ne, 3).sum(axis=(0, 2))
    inv_tot = np.where(tot > 1e-30, 1.0 / np.where(tot > 1e-30, tot, 1.0), 0.0)
    return lam0, sd, inv_tot


class Force3DPlan:
    """Device tables for one (molecule, beam list, Gamma_eff_factor)."""

    def __init__(self, mol_data, beams: Sequence[Beam3D], Gamma_eff_factor=1.0, hbar_value=hbar,
                 basis: str = "auto"):
        if basis not in ("auto", "spin", "jacobi"):
            raise ValueError(f"basis must be 'auto', 'spin' or 'jacobi', not {basis!r}")
        _lib.require_gpu()
        m = mol_data
        ng, ne = m.n_ground, m.n_excited
        Z = np.stack([np.asarray(m.zeeman_x)[:ng, :ng], np.asarray(m.zeeman_y)[:ng, :ng],
                      np.asarray(m.zeeman_z)[:ng, :ng]]).astype(complex)
        zde = np.real(np.diag(np.asarray(m.zeeman_z)[ng:, ng:]))
        dabs = np.abs(np.asarray(m.tdm)[:ng, ng:ng + ne, :]).reshape(ng, 3 * ne)
        nb = len(beams)
        bk = np.array([b.k_hat for b in beams], dtype=float).reshape(nb, 3)
        bf = np.array([b.freq for b in beams], dtype=float)
        bq = np.array([b.q_lab for b in beams], dtype=np.int32)
        bs = np.array([b.s0 for b in beams], dtype=float)
        bw = np.array([b.beam_radius if b.beam_radius is not None else 0.0 for b in beams], dtype=float)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (m.energies, Z.real, Z.imag, zde, dabs, bk, bf)]
        self._keep = arrs + [bq, bs, bw]
        dp = [a.ctypes.data_as(_lib._dp) for a in arrs]
        self._h = ctypes.c_void_p()
        rc = _lib.lib().mm_force3d_plan_create(ctypes.byref(self._h), ng, ne, nb, *dp, bq.ctypes.data_as(_lib._ip),
                                               bs.ctypes.data_as(_lib._dp), bw.ctypes.data_as(_lib._dp),
                                               float(m.Gamma), float(m.k), float(Gamma_eff_factor),
                                               float(hbar_value), float(m.mass))
        _lib.check(rc, "mm_force3d_plan_create")
        self.n_ground = ng
        self.basis = "jacobi"
        tabs = spin_basis_tables(m) if basis != "jacobi" else None
        if basis == "spin" and tabs is None:
            raise ValueError("basis='spin': the ground Zeeman blocks are not spin-1/2 operators")
        if tabs is not None:
            lam0, sd, inv_tot = (np.ascontiguousarray(a, dtype=np.float64) for a in tabs)
            self._keep += [lam0, sd, inv_tot]
            rc = _lib.lib().mm_force3d_plan_set_spin_basis(self._h, lam0.ctypes.data_as(_lib._dp),
                                                           sd.ctypes.data_as(_lib._dp),
                                                           inv_tot.ctypes.data_as(_lib._dp))
            _lib.check(rc, "mm_force3d_plan_set_spin_basis")
            self.basis = "spin"

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            _lib.check(_lib.lib().mm_force3d_plan_destroy(self._h), "mm_force3d_plan_destroy")
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def force3d_points_device(plan: Force3DPlan, r: torch.Tensor, v: torch.Tensor, B_gradient: float, stream=None):
    """(F (n,3), R (n,), pop (n, n_g)) device tensors at r, v (n, 3) [m, m/s]."""
    r = r.contiguous().to(torch.float64)
    v = v.contiguous().to(torch.float64)
    n = r.shape[0]
    F = torch.empty((n, 3), dtype=torch.float64, device=r.device)
    R = torch.empty(n, dtype=torch.float64, device=r.device)
    p = torch.empty((n, plan.n_ground), dtype=torch.float64, device=r.device)
    rc = _lib.lib().mm_force3d_points(plan._h, n, _lib.ptr(r), _lib.ptr(v), float(B_gradient), _lib.ptr(F),
                                      _lib.ptr(R), _lib.ptr(p), _lib.stream_handle(stream))
    _lib.check(rc, "mm_force3d_points")
    return F, R, p


def leapfrog3d_device(plan: Force3DPlan, r0: torch.Tensor, v0: torch.Tensor, B_gradient: float, dt: float,
                      n_steps: int, r_escape: float = 0.015, save_full: bool = True, stream=None,
                      segmented: Optional[bool] = None, seg_steps: int = 0):
    """Leapfrog of n particles; returns dict with r_traj, v_traj
    (n, n_steps, 3) if save_full, r_final, v_half_final (n, 3), n_actual (n,).
    With the spin-rotation basis the run goes in segments (``segmented``,
    default on for that basis): after each segment of ``seg_steps`` (0: 256)
    steps only the particles still inside r_escape continue, compacted, with
    up to 16 lanes each (mm_force3d_leapfrog_ws); otherwise one launch runs
    every particle to the end (mm_force3d_leapfrog)."""
    r0 = r0.contiguous().to(torch.float64)
    v0 = v0.contiguous().to(torch.float64)
    n = r0.shape[0]
    dev = r0.device
    out = {"r_final": torch.empty((n, 3), dtype=torch.float64, device=dev),
           "v_half_final": torch.empty((n, 3), dtype=torch.float64, device=dev),
           "n_actual": torch.empty(n, dtype=torch.int32, device=dev)}
    if save_full:
        out["r_traj"] = torch.empty((n, n_steps, 3), dtype=torch.float64, device=dev)
        out["v_traj"] = torch.empty((n, n_steps, 3), dtype=torch.float64, device=dev)
    if segmented is None:
        segmented = plan.basis == "spin"
    if segmented and n > 0:
        nb = int(_lib.lib().mm_force3d_leapfrog_workspace_bytes(n))
        ws = torch.empty(nb, dtype=torch.uint8, device=dev)
        rc = _lib.lib().mm_force3d_leapfrog_ws(plan._h, n, _lib.ptr(r0), _lib.ptr(v0), float(B_gradient),
                                               float(dt), int(n_steps), float(r_escape), _lib.ptr(out.get("r_traj")),
                                               _lib.ptr(out.get("v_traj")), _lib.ptr(out["r_final"]),
                                               _lib.ptr(out["v_half_final"]), _lib.ptr(out["n_actual"]),
                                               int(seg_steps), _lib.ptr(ws), nb, _lib.stream_handle(stream))
        _lib.check(rc, "mm_force3d_leapfrog_ws")
        return out
    rc = _lib.lib().mm_force3d_leapfrog(plan._h, n, _lib.ptr(r0), _lib.ptr(v0), float(B_gradient), float(dt),
                                        int(n_steps), float(r_escape), _lib.ptr(out.get("r_traj")),
                                        _lib.ptr(out.get("v_traj")), _lib.ptr(out["r_final"]),
                                        _lib.ptr(out["v_half_final"]), _lib.ptr(out["n_actual"]),
                                        _lib.stream_handle(stream))
    _lib.check(rc, "mm_force3d_leapfrog")
    return out


def _dev():
    return torch.device("cuda", torch.cuda.current_device())


def rate_eq_force_3d_batch(mol_data, beams, r_arr, v_arr, B_gradient, Gamma_eff_factor=1.0, plan=None):
    """(F (n,3) N, R (n,) rad/s, pop (n, n_g)) numpy, as
    molmot_cuda/simulator_3d_cuda.py:666-754."""
    plan = plan or Force3DPlan(mol_data, beams, Gamma_eff_factor)
    d = _dev()
    r = torch.from_numpy(np.asarray(r_arr, dtype=np.float64).reshape(-1, 3)).to(d)
    v = torch.from_numpy(np.asarray(v_arr, dtype=np.float64).reshape(-1, 3)).to(d)
    F, R, p = force3d_points_device(plan, r, v, B_gradient)
    return F.cpu().numpy(), R.cpu().numpy(), p.cpu().numpy()


def rate_eq_force_3d(mol_data, beams, r, v, B_gradient, Gamma_eff_factor=1.0, beam_radius=None):
    """(F (3,), R, p (n_g,)) at one point (simulator_3d.py:168-365)."""
    F, R, p = rate_eq_force_3d_batch(mol_data, beams, np.asarray(r, float)[None], np.asarray(v, float)[None],
                                     B_gradient, Gamma_eff_factor)
    return F[0], float(R[0]), p[0]


class MOTSimulator3D:
    """simulator_3d.py:372-532 with device force evaluation."""

    def __init__(self, mol_data, beam_config, B_gradient, s0=1.0, Gamma_eff_factor=1.0, beam_radius=10e-3,
                 basis="auto"):
        self.mol_data = mol_data
        self.basis = basis
        self.B_gradient = B_gradient
        self.Gamma_eff_factor = Gamma_eff_factor
        self.beam_radius = beam_radius
        self.mass = mol_data.mass
        if isinstance(beam_config, list) and len(beam_config) > 0:
            if isinstance(beam_config[0], Beam3D):
                self.beams = beam_config
            elif isinstance(beam_config[0], dict):
                self.beams = make_6beam_config(mol_data, beam_config, beam_radius=beam_radius)
            else:
                raise ValueError("beam_config must be list of Beam3D or beam_pairs dicts")
        else:
            raise ValueError("beam_config must be a non-empty list")
        self._plan = None

    @property
    def plan(self) -> Force3DPlan:
        if self._plan is None:
            self._plan = Force3DPlan(self.mol_data, self.beams, self.Gamma_eff_factor, basis=self.basis)
        return self._plan

    def force_3d(self, r, v):
        F, R, p = rate_eq_force_3d_batch(self.mol_data, self.beams, np.asarray(r, float)[None],
                                         np.asarray(v, float)[None], self.B_gradient, plan=self.plan)
        return F[0], float(R[0]), p[0]

    def force_3d_batch(self, r_arr, v_arr):
        return rate_eq_force_3d_batch(self.mol_data, self.beams, r_arr, v_arr, self.B_gradient, plan=self.plan)

    def simulate_trajectory_3d(self, r0, v0, t_max, dt=1e-6, r_escape=0.015):
        t, r, v = self._ensemble([r0], [v0], t_max, dt, r_escape)
        return t, r[0], v[0]

    def simulate_ensemble_3d(self, r0_list, v0_list, t_max, dt=1e-6, r_escape=0.015, callback=None):
        t, r, v = self._ensemble(r0_list, v0_list, t_max, dt, r_escape)
        out = []
        for i in range(r.shape[0]):
            out.append((t, r[i], v[i]))
            if callback is not None:
                callback(i, t, r[i], v[i])
        return out

    def _ensemble(self, r0_list, v0_list, t_max, dt, r_escape):
        n_steps = int(t_max / dt)
        d = _dev()
        r0 = torch.from_numpy(np.asarray(r0_list, dtype=np.float64).reshape(-1, 3)).to(d)
        v0 = torch.from_numpy(np.asarray(v0_list, dtype=np.float64).reshape(-1, 3)).to(d)
        out = leapfrog3d_device(self.plan, r0, v0, self.B_gradient, dt, n_steps, r_escape)
        return np.arange(n_steps) * dt, out["r_traj"].cpu().numpy(), out["v_traj"].cpu().numpy()


def compute_3d_force_map(simulator, axes="xz", n_points=50, v=None, extent=5e-3):
    """(c1, c2, F_a1 (n,n), F_a2 (n,n), R (n,n)) on a plane through the trap
    (simulator_3d.py:539-595), all points in one launch."""
    v = np.zeros(3) if v is None else np.asarray(v, float)
    ax = {"x": 0, "y": 1, "z": 2}
    a1, a2 = ax[axes[0]], ax[axes[1]]
    c1 = np.linspace(-extent, extent, n_points)
    c2 = np.linspace(-extent, extent, n_points)
    r = np.zeros((n_points, n_points, 3))
    r[:, :, a1] = c1[:, None]
    r[:, :, a2] = c2[None, :]
    F, R, _ = simulator.force_3d_batch(r.reshape(-1, 3), np.tile(v, (n_points * n_points, 1)))
    F = F.reshape(n_points, n_points, 3)
    return c1, c2, F[:, :, a1], F[:, :, a2], R.reshape(n_points, n_points)


def compute_3d_force_map_hip(mol_data, beams, axes="xz", n_points=50, v=None, B_gradient=0.0, extent=5e-3,
                             Gamma_eff_factor=1.0):
    """compute_3d_force_map_cuda (simulator_3d_cuda.py:757-815)."""
    sim = MOTSimulator3D(mol_data, list(beams), B_gradient, Gamma_eff_factor=Gamma_eff_factor)
    return compute_3d_force_map(sim, axes, n_points, v, extent)


def simulate_trajectory_3d_hip(mol_data, beams, r0, v0, t_max, dt=1e-6, B_gradient=0.0, Gamma_eff_factor=1.0,
                               r_escape=0.015):
    """simulate_trajectory_3d_cuda (simulator_3d_cuda.py:818-887)."""
    return MOTSimulator3D(mol_data, list(beams), B_gradient, Gamma_eff_factor=Gamma_eff_factor
                          ).simulate_trajectory_3d(r0, v0, t_max, dt, r_escape)


def simulate_ensemble_3d_hip(mol_data, beams, r0_list, v0_list, t_max, dt=1e-6, B_gradient=0.0,
                             Gamma_eff_factor=1.0, r_escape=0.015):
    """simulate_ensemble_3d_cuda (simulator_3d_cuda.py:890-973): (t, r_all
    (n, n_steps, 3), v_all (n, n_steps, 3))."""
    sim = MOTSimulator3D(mol_data, list(beams), B_gradient, Gamma_eff_factor=Gamma_eff_factor)
    return sim._ensemble(r0_list, v0_list, t_max, dt, r_escape)


compute_3d_force_map_cuda = compute_3d_force_map_hip
simulate_trajectory_3d_cuda = simulate_trajectory_3d_hip
simulate_ensemble_3d_cuda = simulate_ensemble_3d_hip
