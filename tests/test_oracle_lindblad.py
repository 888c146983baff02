"""Pin the Lindblad oracle (CPU): two-level optical Bloch closed form,
Liouvillian invariants, and the RK4 propagator's long-time limit against the
steady-state solve of molmot/obe/obe_subdoppler.py."""
import numpy as np
import pytest

from oracle import lindblad as lb
from molmot_amd.molecules.sroh import MolecularData, GAMMA_RAD


def two_level(E_e=4.0e14):
    tdm = np.zeros((2, 2, 3), dtype=complex)
    tdm[0, 1, 2] = 1.0
    tdm[1, 0, 2] = 1.0
    Z = np.zeros((2, 2), dtype=complex)
    return MolecularData(energies=np.array([0.0, E_e]), tdm=tdm, zeeman_x=Z, zeeman_y=Z, zeeman_z=Z,
                         n_ground=1, n_excited=1, n_states=2)


@pytest.mark.parametrize("det_G,Om_G", [(0.0, 0.5), (-1.0, 1.0), (2.0, 3.0)])
def test_two_level_closed_form(det_G, Om_G):
    """rho_ee = (Omega^2/4) / (delta^2 + Gamma^2/4 + Omega^2/2) for one
    travelling wave; F = hbar k Gamma rho_ee along +k."""
    m = two_level()
    G = m.Gamma
    delta_hz = det_G * G / (2 * np.pi)
    pairs = [{"freq": m.energies[1] + delta_hz, "q_fwd": 2, "q_bwd": 0, "s0": 1.0}]
    F, rho, R = lb.force_subdoppler(m, pairs, 0.0, 0.0, 0.0, Om_G, n_spatial=4)
    Om = Om_G * G
    d = det_G * G
    pe = (Om ** 2 / 4) / (d ** 2 + G ** 2 / 4 + Om ** 2 / 2)
    assert abs(rho[1, 1].real - pe) < 1e-12
    assert abs(F - lb.HBAR * m.k * G * pe) < 1e-10 * lb.HBAR * m.k * G
    assert abs(R - G * pe) < 1e-9 * G


def test_liouvillian_preserves_trace(mol):
    pairs = [{"freq": mol.omega_J32 - 2e6, "q_fwd": 2, "q_bwd": 0, "s0": 1.0},
             {"freq": mol.omega_J12 - 2e6, "q_fwd": 0, "q_bwd": 2, "s0": 1.0}]
    H = lb.hamiltonian(mol, pairs, 0.0, 2e-3, 0.7, 0.2, 1.0)
    assert np.allclose(H, H.conj().T)
    L = lb.liouvillian(H, mol, mol.Gamma)
    N = mol.n_states
    tr = np.zeros(N * N)
    tr[[i * N + i for i in range(N)]] = 1.0
    assert np.max(np.abs(tr @ L)) < 1e-6 * np.max(np.abs(L))


def test_rk4_long_time_equals_steady_state_two_level():
    m = two_level()
    G = m.Gamma
    pairs = [{"freq": m.energies[1] - 0.7 * G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0}]
    F_ss, _, R_ss = lb.force_subdoppler(m, pairs, 0.0, 0.0, 0.0, 1.2, n_spatial=1)
    F_rk, R_rk, rho = lb.rk4_force(m, pairs, 0.0, 0.0, 0.0, 0.0, 1.2, t_total=60.0, dt=0.02)
    assert abs(F_rk - F_ss) < 1e-9 * abs(F_ss)
    assert abs(np.trace(rho) - 1) < 1e-12


def test_rk4_long_time_equals_steady_state_multilevel():
    """Synthetic 2+2 system with two beam pairs (one sigma, one pi) at
    B != 0: the Liouvillian has a unique steady state and a spectral gap of
    ~0.16 Gamma, so RK4 from equal ground populations reaches the steady state
    of obe_subdoppler.py:144-183 well within t = 400 / Gamma.  (For SrOH with
    several pairs the gap is 1e-3..1e-5 Gamma because of near-dark ground
    superpositions, too slow for this check.)"""
    from molmot_amd.molecules import synthetic_molecule
    m = synthetic_molecule(2, 2, seed=4)
    G = m.Gamma
    Ee = np.mean(m.energies[2:])
    pairs = [{"freq": Ee - m.energies[0] - 0.8 * G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0},
             {"freq": Ee - m.energies[1] - 0.5 * G / (2 * np.pi), "q_fwd": 1, "q_bwd": 1, "s0": 1.0}]
    phi = 0.4
    H = lb.hamiltonian(m, pairs, 0.0, 2e-3, phi, 0.5, 1.0)
    L = lb.liouvillian(H, m, G)
    w = np.sort(np.linalg.eigvals(L / G).real)[::-1]
    assert w[1] < -0.1
    rho_ss = lb.steady_state(L, m.n_states)
    A, B = lb.coupling_blocks(m, pairs, 1.0)
    F_ss = lb.force_from_rho(rho_ss, m, A * np.exp(1j * phi), B * np.exp(-1j * phi))
    F_rk, R_rk, rho = lb.rk4_force(m, pairs, 0.0, phi, 2e-3, 0.5, 1.0, t_total=400.0, dt=0.02, avg_from=0.9)
    assert np.max(np.abs(rho - rho_ss)) < 1e-10
    assert abs(F_rk - F_ss) < 1e-9 * abs(F_ss)
