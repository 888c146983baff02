"""Dipole-force expectation values (molmot/obe/force.py:23-197).

Natural units (hbar k Gamma) throughout; E_kq is the (6, 3) six-beam field
of compute_fields_6beam (rows +x +y +z -x -y -z)."""
from __future__ import annotations

import numpy as np


def _force_from_dexp(d_exp, E_kq):
    """F_k = -2 sum_q Im(d_exp[q] (E_kq[k,q] - E_kq[k+3,q]))."""
    Ed = E_kq[:3] - E_kq[3:]
    return -2.0 * np.imag(Ed @ d_exp)


def force_from_wavefunction(psi, d_ge, E_kq, n_ground, n_excited, k_dirs=None):
    """<psi|F|psi> for amplitudes psi (interaction picture) (force.py:23-99):
    d_exp[q] = sum_g conj(psi_g) sum_e d_ge[g,e,q] psi_e."""
    pg = psi[:n_ground]
    pe = psi[n_ground:n_ground + n_excited]
    d_exp = np.conj(pg) @ np.einsum("geq,e->gq", d_ge, pe)
    return _force_from_dexp(d_exp, E_kq)


def force_from_density_matrix(rho, d_ge, E_kq, n_ground, n_excited):
    """Same bilinear form with conj(psi_g) psi_e -> rho[e, g] (force.py:102-146)."""
    d_exp = np.einsum("geq,eg->q", d_ge, rho[n_ground:n_ground + n_excited, :n_ground])
    return _force_from_dexp(d_exp, E_kq)


def zeeman_gradient_force(psi, zeeman_matrices, B_gradient, mu_B_Gamma):
    """-mu_B dB_i/dr_i <psi|Z_i|psi> for the quadrupole gradients
    (B'/2, B'/2, -B') (force.py:149-197)."""
    grad = np.array([B_gradient / 2, B_gradient / 2, -B_gradient])
    F = np.zeros(3)
    for i, ax in enumerate("xyz"):
        F[i] = -mu_B_Gamma * grad[i] * np.real(np.conj(psi) @ zeeman_matrices[ax] @ psi)
    return F
