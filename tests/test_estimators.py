"""The ensemble estimators of molmot/obe/diffusion.py:35-239 and the dipole
force of molmot/obe/force.py:23-197 (molmot_amd.obe.diffusion / .force),
CPU only: closed forms on constructed ensembles, statistical recovery of a
set diffusion constant and damping rate, the device (torch) variants equal
to the list variants, and the force against the reference's own _rhs
vectors (tests/golden/ref_sse.npz)."""
import os

import numpy as np
import pytest
import torch

from molmot_amd.obe import diffusion as dif
from molmot_amd.obe import force as frc
from molmot_amd.obe.stochastic import SSEResult

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MASS = 105 * 1.66053906660e-27


def _results(times, V, jumps=None, photons=None):
    out = []
    for i in range(V.shape[0]):
        out.append(SSEResult(times=times, positions=np.zeros_like(V[i]), velocities=V[i],
                             populations=np.zeros((times.size, 1)),
                             photons_scattered=0 if photons is None else int(photons[i]),
                             jump_times=[] if jumps is None else list(jumps[i])))
    return out


def test_diffusion_closed_form():
    """v = +-sqrt(a t) on two trajectories: mean 0 and Var v = a t exactly,
    so D = m^2 a / 2 and the residual-based error is 0."""
    a = 3.7e-3
    t = np.linspace(0.0, 2e-3, 41)
    V = np.zeros((2, t.size, 3))
    V[0, :, 2] = np.sqrt(a * t)
    V[1, :, 2] = -np.sqrt(a * t)
    D, Derr, T = dif.compute_diffusion_from_ensemble(_results(t, V), MASS)
    assert abs(D - MASS ** 2 * a / 2) < 1e-12 * MASS ** 2 * a / 2
    assert Derr < 1e-9 * D and np.isnan(T)
    Dd = dif.compute_diffusion_from_ensemble_device(torch.from_numpy(V), t, MASS)
    assert abs(Dd[0] - D) < 1e-12 * D


def test_diffusion_recovers_brownian_constant():
    """Brownian velocities with dVar(v)/dt = 2 D / m^2: D within 4 standard
    errors of the ensemble (N = 20,000)."""
    rng = np.random.default_rng(5)
    D0 = 2.0e-49                      # kg^2 m^2 / s^3
    t = np.linspace(0.0, 1e-3, 26)
    dt = t[1] - t[0]
    s = np.sqrt(2 * D0 / MASS ** 2 * dt)
    inc = rng.normal(0.0, s, (20000, t.size - 1))
    V = np.zeros((20000, t.size, 3))
    V[:, 1:, 0] = np.cumsum(inc, axis=1)
    D, Derr, _ = dif.compute_diffusion_from_ensemble(_results(t, V), MASS, axis=0)
    # relative sampling error of a variance from N samples ~ sqrt(2/N) = 1%
    assert abs(D - D0) < 0.04 * D0
    Dd = dif.compute_diffusion_from_ensemble_device(torch.from_numpy(V), t, MASS, axis=0)
    assert abs(Dd[0] - D) < 1e-9 * D


def test_damping_closed_form_and_statistics():
    """v_i(t) = v0_i exp(-beta t): the normalised autocorrelation is exactly
    exp(-beta t), so beta comes back to rounding; an Ornstein-Uhlenbeck
    ensemble in equilibrium gives beta within 3%."""
    beta = 850.0
    t = np.linspace(0.0, 3e-3, 31)
    v0 = np.array([0.3, -1.2, 0.7, 2.0])
    V = np.zeros((4, t.size, 3))
    V[:, :, 2] = v0[:, None] * np.exp(-beta * t)[None, :]
    assert abs(dif.estimate_damping_rate(_results(t, V), MASS) - beta) < 1e-9 * beta
    assert abs(dif.estimate_damping_rate_device(torch.from_numpy(V), t) - beta) < 1e-9 * beta
    rng = np.random.default_rng(9)
    n = 40000
    sig = 0.5
    V = np.zeros((n, t.size, 3))
    V[:, 0, 1] = rng.normal(0, sig, n)
    dt = t[1] - t[0]
    f = np.exp(-beta * dt)
    for j in range(1, t.size):
        V[:, j, 1] = f * V[:, j - 1, 1] + rng.normal(0, sig * np.sqrt(1 - f * f), n)
    b = dif.estimate_damping_rate(_results(t, V), MASS, axis=1)
    assert abs(b - beta) < 0.03 * beta
    # a signal that never decays below 0.1 ... and one with no variance
    assert dif.estimate_damping_rate(_results(t, np.zeros((3, t.size, 3))), MASS) == 0.0


def test_diffusion_temperature():
    assert dif.compute_diffusion_temperature(1.0, 2.0, 0.0) == float("inf")
    T = dif.compute_diffusion_temperature(3e-49, MASS, 900.0)
    assert abs(T - 3e-49 / (MASS * 900.0 * 1.380649e-23)) < 1e-15 * T


def test_scattering_rate():
    t = np.linspace(0.0, 1e-3, 11)
    V = np.zeros((3, t.size, 3))
    jumps = [[1e-4, 2e-4, 7e-4], [5e-4], []]
    R, Rerr = dif.compute_scattering_rate(_results(t, V, jumps))
    rates = np.array([3, 1, 0]) / 1e-3
    assert abs(R - rates.mean()) < 1e-9 * R and abs(Rerr - rates.std() / np.sqrt(3)) < 1e-9 * Rerr
    R2, _ = dif.compute_scattering_rate(_results(t, V, jumps), t_min=4e-4)
    assert abs(R2 - np.mean(np.array([1, 1, 0]) / 6e-4)) < 1e-9 * R2
    # photon counts only (the GPU ensemble without a jump log)
    R3, _ = dif.compute_scattering_rate(_results(t, V, None, [3, 1, 0]))
    assert abs(R3 - R) < 1e-9 * R
    with pytest.raises(ValueError, match="log_jumps"):
        dif.compute_scattering_rate(_results(t, V, None, [3, 1, 0]), t_min=4e-4)


def test_force_from_wavefunction_vs_reference_rhs():
    """The force of SSESolver._rhs (stochastic.py:862-878) written by the
    reference into ref_sse.npz: F = force_from_wavefunction(psi exp(-i w t),
    d_ge, E_kq) / m_nat, and force_from_density_matrix(|psi><psi|) agrees."""
    from oracle import mcwf as mc
    from molmot_amd.molecules import load_sroh_from_julia
    d = np.load(os.path.join(GOLDEN, "ref_sse.npz"), allow_pickle=False)
    mol = load_sroh_from_julia()
    pairs = [dict(freq=float(f), q_fwd=int(a), q_bwd=int(b), s0=float(s))
             for f, a, b, s in zip(d["pairs_freq"], d["pairs_q_fwd"], d["pairs_q_bwd"], d["pairs_s0"])]
    tab = mc.tables(mol, pairs, float(d["B_gradient"]), beam_radius=float(d["beam_radius"]))
    ng, ne, N = tab["ng"], tab["ne"], tab["N"]
    for i in range(d["t"].size):
        pint = d["psi"][i] * np.exp(-1j * d["omega0s"][i] * d["t"][i])
        F = frc.force_from_wavefunction(pint, tab["d_ge"], d["fields"][i], ng, ne)
        ref = d["du"][i][2 * N + 3:] * tab["mass_nat"]
        assert np.max(np.abs(F - ref)) < 1e-12 * np.max(np.abs(ref)), i
        rho = np.outer(pint, np.conj(pint))
        Fr = frc.force_from_density_matrix(rho, tab["d_ge"], d["fields"][i], ng, ne)
        assert np.max(np.abs(Fr - F)) < 1e-13 * np.max(np.abs(F)), i


def test_zeeman_gradient_force_closed_form():
    """For basis state |i> and diagonal Z_a: F_a = -mu dB_a/dr_a Z_a[i, i]
    with the quadrupole gradients (B'/2, B'/2, -B') (force.py:149-197)."""
    Z = {a: np.diag(np.arange(1.0, 5.0) * (k + 1)).astype(complex) for k, a in enumerate("xyz")}
    psi = np.zeros(4, complex)
    psi[2] = 1.0
    F = frc.zeeman_gradient_force(psi, Z, 0.2, 1.5)
    assert np.allclose(F, [-1.5 * 0.1 * 3.0, -1.5 * 0.1 * 6.0, 1.5 * 0.2 * 9.0], rtol=1e-15, atol=0)
