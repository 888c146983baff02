"""Floquet OBE sub-Doppler force on MI355X (molmot/obe/floquet.py:61-403,
molmot_cuda/floquet_cuda.py:230-478).

For a molecule moving at v through the standing wave the Lindblad generator
is periodic, L(t) = L0 + L_plus e^{i kv t} + L_minus e^{-i kv t}, and the
Fourier components of its periodic steady state solve the block-tridiagonal
system (L0 - i n kv) rho_n + L_plus rho_{n-1} + L_minus rho_{n+1} = 0,
n = -n_max .. n_max, with the last row of the n = 0 block replaced by
Tr rho_0 = 1 (floquet.py:194-300).  The reference factorises the whole
(2 n_max + 1) N^2 system densely per velocity.  Here every velocity of a
batch is solved by block elimination from both ends towards n = 0:

    top:    G_{-n_max} = D_{-n_max},  X_b = G_b^{-1} L_minus,
            G_{b+1} = D_{b+1} - L_plus X_b            (b = -n_max .. -1)
    bottom: H_{n_max} = D_{n_max},   Z_b = H_b^{-1} L_plus,
            H_{b-1} = D_{b-1} - L_minus Z_b           (b = n_max .. 1)
    centre: (D_0 - L_plus X_{-1} - L_minus Z_1) rho_0 = e_trace
    then    rho_{-1} = -X_{-1} rho_0,  rho_{+1} = -Z_1 rho_0

with D_b = L0 - i b kv.  The N^2 x N^2 solves (N^2 right-hand sides) run in
the library's batched LU kernel (mm_zgesv_batched, one workgroup per
velocity); the Schur updates are plain batched complex GEMMs.  Only rho_0 and
rho_{+-1} are formed: the force needs no other component.  Velocities whose
elimination meets an exactly singular block are re-solved densely by least
squares, the reference's fallback (floquet.py:291-293).

coupling="reference" keeps the reference's model (H_plus, H_minus each
Hermitian); coupling="standing_wave" uses the Hamiltonian of
obe_subdoppler.build_H_standing_wave, whose periodic steady state the RK4
propagator (molmot_amd.obe.lindblad) reaches in the time domain.
"""
from __future__ import annotations

import numpy as np
import torch

from ..constants import hbar as HBAR
from ..hip import _lib


def _dissipator(mol, Gamma_eff, N):
    """Lindblad dissipator of the jump operators sqrt(Gamma_eff br)|g><e|
    with br >= 1e-15 (floquet.py:161-181), column-major vec."""
    ng, ne = mol.n_ground, mol.n_excited
    d2 = mol.d_squared
    L = np.zeros((N * N, N * N), dtype=complex)
    for ie in range(ne):
        e = ng + ie
        total = float(np.sum(d2[:, ie, :]))
        if total < 1e-30:
            continue
        for ig in range(ng):
            rate = 0.0
            for q in range(3):
                br = d2[ig, ie, q] / total
                if br >= 1e-15:
                    rate += Gamma_eff * br
            if rate == 0.0:
                continue
            # repopulation rho_ee -> rho_gg; decay of the e row and column
            L[ig + N * ig, e + N * e] += rate
            for j in range(N):
                L[e + N * j, e + N * j] -= 0.5 * rate
                L[j + N * e, j + N * e] -= 0.5 * rate
    return L


def _commutator(H):
    """Superoperator of -i [H, .] in the column-major vec."""
    N = H.shape[0]
    I = np.eye(N, dtype=complex)
    return -1j * (np.kron(I, H) - np.kron(H.T, I))


class FloquetOBE:
    """FloquetOBE(mol_data, beam_pairs, B_gradient, Omega_Gamma=1.0, n_max=3,
    Gamma_eff_factor=1.0) with the reference's methods (floquet.py:61-403);
    velocities are solved as one batch on the GPU."""

    def __init__(self, mol_data, beam_pairs, B_gradient, Omega_Gamma=1.0, n_max=3, Gamma_eff_factor=1.0,
                 coupling: str = "reference"):
        if coupling not in ("reference", "standing_wave"):
            raise ValueError("coupling is 'reference' or 'standing_wave'")
        if n_max < 0:
            raise ValueError("n_max >= 0 required")
        self.mol_data = mol_data
        self.beam_pairs = beam_pairs
        self.B_gradient = B_gradient
        self.Omega_Gamma = Omega_Gamma
        self.n_max = int(n_max)
        self.Gamma_eff_factor = Gamma_eff_factor
        self.coupling = coupling
        self.Gamma = mol_data.Gamma
        self.k = mol_data.k
        self.n_g, self.n_e = mol_data.n_ground, mol_data.n_excited
        self.N = self.n_g + self.n_e
        self.N2 = self.N * self.N
        self._build_floquet_blocks()
        self._dev = None

    def _build_floquet_blocks(self):
        """H0, H_plus, H_minus and the Liouvillian blocks (floquet.py:78-192)."""
        mol, N, ng, ne = self.mol_data, self.N, self.n_g, self.n_e
        E = np.asarray(mol.energies, dtype=float)
        Ee = np.mean(E[ng:])
        lf = [bp["freq"] for bp in self.beam_pairs]
        H0 = np.zeros((N, N), dtype=complex)
        for ig in range(ng):
            om = Ee - E[ig]
            H0[ig, ig] = -2 * np.pi * min([f - om for f in lf], key=abs)
        for ie in range(ne):
            H0[ng + ie, ng + ie] = 2 * np.pi * (E[ng + ie] - Ee)
        Hp = np.zeros((N, N), dtype=complex)
        Hm = np.zeros((N, N), dtype=complex)
        A = np.zeros((N, N), dtype=complex)
        B = np.zeros((N, N), dtype=complex)
        Om0 = self.Omega_Gamma * self.Gamma
        d = mol.tdm
        ref = self.coupling == "reference"
        for bp in self.beam_pairs:
            Om = Om0 * np.sqrt(bp["s0"])
            for ig in range(ng):
                for ie in range(ne):
                    e = ng + ie
                    df = abs(d[ig, e, bp["q_fwd"]])
                    if df > 1e-10:
                        c = Om * df / 2
                        A[e, ig] += c
                        Hp[e, ig] += c
                        (Hp if ref else Hm)[ig, e] += np.conj(c)
                    db = abs(d[ig, e, bp["q_bwd"]])
                    if db > 1e-10:
                        c = Om * db / 2
                        B[e, ig] += c
                        Hm[e, ig] += c
                        (Hm if ref else Hp)[ig, e] += np.conj(c)
        self._H0, self._H_plus, self._H_minus, self._A, self._B = H0, Hp, Hm, A, B
        self._L0 = _commutator(H0) + _dissipator(mol, self.Gamma * self.Gamma_eff_factor, N)
        self._L_plus = _commutator(Hp)
        self._L_minus = _commutator(Hm)

    def _device(self):
        if self._dev is None:
            dev = torch.device("cuda", torch.cuda.current_device())
            self._dev = {k: torch.from_numpy(np.ascontiguousarray(getattr(self, "_" + k))).to(dev)
                         for k in ("L0", "L_plus", "L_minus", "H_plus", "H_minus", "A", "B")}
        return self._dev

    def _zgesv(self, A, B, stream=None):
        """A X = B in place for a batch (A (nv, n, n), B (nv, n, m)); returns
        (X, info)."""
        nv, n = A.shape[0], A.shape[1]
        info = torch.zeros(nv, dtype=torch.int32, device=A.device)
        rc = _lib.lib().mm_zgesv_batched(n, B.shape[2], nv, _lib.ptr(A), _lib.ptr(B), _lib.ptr(info), None, None,
                                         _lib.stream_handle(stream))
        _lib.check(rc, "mm_zgesv_batched")
        return B, info

    def solve_components(self, v_arr, z_macro: float = 0.0, stream=None):
        """(rho_0, rho_-1, rho_+1), each (nv, N, N) complex128 on the device,
        for the velocities v_arr (exact zeros replaced by 1e-6 m/s as the
        reference does, floquet.py:334-340)."""
        _lib.require_gpu()
        t = self._device()
        N, N2, nmax = self.N, self.N2, self.n_max
        v = np.asarray(v_arr, dtype=np.float64).ravel()
        v = np.where(np.abs(v) < 1e-12, 1e-6, v)
        nv = v.size
        dev = t["L0"].device
        L0 = t["L0"]
        if abs(z_macro) > 1e-15:
            Bg = self.B_gradient * z_macro * 1e4
            EZ = 2 * np.pi * Bg * np.asarray(self.mol_data.zeeman_z_diag) * self.Gamma / (2 * np.pi)
            L0 = L0 + torch.from_numpy(_commutator(np.diag(EZ).astype(complex))).to(dev)
        kv = torch.from_numpy(self.k * v).to(dev).to(torch.complex128)
        eye = torch.arange(N2, device=dev)

        def D(n):
            M = L0.expand(nv, N2, N2).clone()
            M[:, eye, eye] -= 1j * n * kv[:, None]
            return M

        Lp, Lm = t["L_plus"], t["L_minus"]
        X = Z = None
        info = torch.zeros(nv, dtype=torch.int32, device=dev)
        for b in range(-nmax, 0):            # top blocks towards the centre
            G = D(b)
            if X is not None:
                G -= torch.matmul(Lp, X)
            X, inf = self._zgesv(G, Lm.expand(nv, N2, N2).clone(), stream)
            info |= inf
        for b in range(nmax, 0, -1):         # bottom blocks towards the centre
            H = D(b)
            if Z is not None:
                H -= torch.matmul(Lm, Z)
            Z, inf = self._zgesv(H, Lp.expand(nv, N2, N2).clone(), stream)
            info |= inf
        S = D(0)
        if X is not None:
            S -= torch.matmul(Lp, X)
        if Z is not None:
            S -= torch.matmul(Lm, Z)
        S[:, N2 - 1, :] = 0.0                 # Tr rho_0 = 1 replaces the last row
        S[:, N2 - 1, eye[::N + 1]] = 1.0
        rhs = torch.zeros((nv, N2, 1), dtype=torch.complex128, device=dev)
        rhs[:, N2 - 1, 0] = 1.0
        x0, inf = self._zgesv(S, rhs, stream)
        info |= inf
        xm = -torch.matmul(X, x0) if X is not None else torch.zeros_like(x0)
        xp = -torch.matmul(Z, x0) if Z is not None else torch.zeros_like(x0)
        bad = torch.nonzero(info != 0).flatten()
        if bad.numel():
            for i in bad.tolist():
                comps = self._dense_lstsq(float(v[i]), L0)
                x0[i, :, 0], xm[i, :, 0], xp[i, :, 0] = comps

        def mat(x):   # column-major vec -> rho[i, j]
            return x[..., 0].reshape(nv, N, N).transpose(1, 2)

        return mat(x0), mat(xm), mat(xp)

    def _dense_lstsq(self, v, L0):
        """The reference's dense system solved by least squares (its fallback)."""
        N2, nmax = self.N2, self.n_max
        nb = 2 * nmax + 1
        dev = L0.device
        t = self._device()
        kv = self.k * v
        M = torch.zeros((nb * N2, nb * N2), dtype=torch.complex128, device=dev)
        I = torch.eye(N2, dtype=torch.complex128, device=dev)
        for b in range(nb):
            n = b - nmax
            r = slice(b * N2, (b + 1) * N2)
            M[r, r] = L0 - 1j * n * kv * I
            if b > 0:
                M[r, (b - 1) * N2:b * N2] = t["L_plus"]
            if b < nb - 1:
                M[r, (b + 1) * N2:(b + 2) * N2] = t["L_minus"]
        row = nmax * N2 + N2 - 1
        M[row, :] = 0.0
        M[row, nmax * N2 + torch.arange(self.N, device=dev) * (self.N + 1)] = 1.0
        rhs = torch.zeros((nb * N2, 1), dtype=torch.complex128, device=dev)
        rhs[row, 0] = 1.0
        x = (torch.linalg.pinv(M) @ rhs)[:, 0]
        return (x[nmax * N2:(nmax + 1) * N2], x[(nmax - 1) * N2:nmax * N2] if nmax else 0 * x[:N2],
                x[(nmax + 1) * N2:(nmax + 2) * N2] if nmax else 0 * x[:N2])

    def _forces(self, rm1, rp1):
        t = self._device()
        if self.coupling == "reference":
            Fp = torch.einsum("vij,ji->v", rm1, t["H_plus"])
            Fm = torch.einsum("vij,ji->v", rp1, t["H_minus"])
            return HBAR * self.k * (Fp.imag - Fm.imag)
        Fa = torch.einsum("vij,ji->v", rm1, t["A"])
        Fb = torch.einsum("vij,ji->v", rp1, t["B"])
        return 2 * HBAR * self.k * (Fa.imag - Fb.imag)

    def force_vs_velocity(self, v_arr, z_macro: float = 0.0, n_points_z: int = 1) -> np.ndarray:
        """Force (N) at each velocity (floquet.py:302-378)."""
        r0, rm1, rp1 = self.solve_components(v_arr, z_macro)
        return self._forces(rm1, rp1).cpu().numpy()

    def scattering_rate(self, v, z_macro: float = 0.0) -> float:
        """Gamma_eff times the excited population of rho_0 (floquet.py:380-403)."""
        r0 = self.solve_components([v], z_macro)[0][0]
        R = self.Gamma * self.Gamma_eff_factor * float(torch.diagonal(r0).real[self.n_g:].sum())
        return max(R, 0.0)

    def scattering_rates(self, v_arr, z_macro: float = 0.0) -> np.ndarray:
        r0 = self.solve_components(v_arr, z_macro)[0]
        R = self.Gamma * self.Gamma_eff_factor * torch.diagonal(r0, dim1=1, dim2=2).real[:, self.n_g:].sum(1)
        return np.maximum(R.cpu().numpy(), 0.0)

    def steady_state_rho(self, v, z_macro: float = 0.0) -> np.ndarray:
        """rho_0 at v (floquet.py:405-418)."""
        return self.solve_components([v], z_macro)[0][0].cpu().numpy()


def force_vs_velocity_cuda(mol_data, beam_pairs, v_arr, B_gradient=0.0, Omega_Gamma=1.0, n_max=3,
                           Gamma_eff_factor=1.0):
    """(F (n_v,), rho0 (n_v, N, N)) (molmot_cuda/floquet_cuda.py:230-333;
    z_macro = 0 as there)."""
    fo = FloquetOBE(mol_data, beam_pairs, B_gradient, Omega_Gamma, n_max, Gamma_eff_factor)
    r0, rm1, rp1 = fo.solve_components(v_arr)
    return fo._forces(rm1, rp1).cpu().numpy(), r0.cpu().numpy()


def force_vs_velocity_cuda_fast(mol_data, beam_pairs, v_arr, B_gradient=0.0, Omega_Gamma=1.0, n_max=3,
                                Gamma_eff_factor=1.0):
    """Forces only (molmot_cuda/floquet_cuda.py:336-420)."""
    return force_vs_velocity_cuda(mol_data, beam_pairs, v_arr, B_gradient, Omega_Gamma, n_max,
                                  Gamma_eff_factor)[0]


def scattering_rate_vs_velocity_cuda(mol_data, beam_pairs, v_arr, B_gradient=0.0, Omega_Gamma=1.0, n_max=3,
                                     Gamma_eff_factor=1.0):
    """Scattering rates (rad/s) (molmot_cuda/floquet_cuda.py:423-478)."""
    fo = FloquetOBE(mol_data, beam_pairs, B_gradient, Omega_Gamma, n_max, Gamma_eff_factor)
    return fo.scattering_rates(v_arr)
