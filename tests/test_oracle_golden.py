"""Pin the CPU oracle against the reference's committed fixtures, closed
forms and invariants (CPU only).

Fixtures (copied from the reference's validation/ folder, data only):
  three_way_results.npz            Hallas red MOT, 101 z and 101 v points:
                                   *_cpu_s  = molmot NumPy solver (CODATA hbar)
                                   *_gpu    = molmot CUDA kernel (legacy hbar)
                                   *_cpu_u  = unsaturated model (== Julia)
  julia_redmot_results.npz         the Julia (OBE.jl) unsaturated result
  trajectory_cpu_vs_gpu_results.npz  4 reference trajectories and a
                                   100-particle batch, 30k Euler steps (JIT path)
Forces in three_way are in units of hbar_CODATA * k * Gamma / 2
(validation/sroh_redmot_three_way.py:196).
"""
import numpy as np
import pytest

import oracle
from molmot_amd.constants import HBAR_LEGACY, hbar as HBAR
from molmot_amd.molecules import synthetic_molecule
from molmot_amd.obe.fields import (LaserBeam, beam_table, make_dc_mot_beams, make_hallas_beams)

B_HALLAS = 21.4 * 1e-4 / 1e-2  # validation/sroh_redmot_three_way.py:60


def maxrel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def hallas(mol):
    beams, freqs, sats = make_hallas_beams(mol)
    return beam_table(beams), freqs, sats


def test_hallas_beam_builder_matches_fixture(mol, golden, hallas):
    jl = golden("julia_redmot_results.npz")
    py = golden("py_redmot_results.npz")
    _, freqs, sats = hallas
    assert np.array_equal(np.asarray(freqs), jl["freqs"])
    assert np.array_equal(np.asarray(freqs), py["freqs_Hz"])
    assert maxrel(sats, py["sats"]) < 1e-15


@pytest.mark.parametrize("flavour,hbar,key", [("numpy", HBAR, "cpu_s"), ("jit", HBAR_LEGACY, "gpu")])
def test_oracle_reproduces_three_way(mol, golden, hallas, flavour, hbar, key):
    tw = golden("three_way_results.npz")
    bt = hallas[0]
    unit = HBAR * mol.k * mol.Gamma / 2.0
    z, v = tw["z_arr"], tw["v_arr"]
    Fz, Pz, _ = oracle.rate_eq_points(mol, bt, np.zeros_like(z), z, B_HALLAS, hbar=hbar, flavour=flavour)
    Fv, _, _ = oracle.rate_eq_points(mol, bt, v, np.zeros_like(v), B_HALLAS, hbar=hbar, flavour=flavour)
    assert maxrel(Fz / unit, tw["Fz_" + key]) < 1e-13
    assert maxrel(Fv / unit, tw["Fv_" + key]) < 1e-13
    assert np.max(np.abs(Pz - tw["Pz_" + key])) < 1e-14
    dz, dv = 0.3e-3, 0.1
    F, _, _ = oracle.rate_eq_points(mol, bt, [0.0, 0.0, dv, -dv], [dz, -dz, 0.0, 0.0], B_HALLAS,
                                    hbar=hbar, flavour=flavour)
    k = -(F[0] - F[1]) / (2 * dz)
    beta = -(F[2] - F[3]) / (2 * dv * mol.mass)
    assert abs(k / tw["k_" + key] - 1) < 1e-12
    assert abs(beta / tw["beta_" + key] - 1) < 1e-12


def test_oracle_unsaturated_matches_julia(mol, golden, hallas):
    tw = golden("three_way_results.npz")
    jl = golden("julia_redmot_results.npz")
    unit = HBAR * mol.k * mol.Gamma / 2.0
    z = tw["z_arr"]
    F, P, _ = oracle.rate_eq_points(mol, hallas[0], np.zeros_like(z), z, B_HALLAS, hbar=HBAR,
                                    flavour="numpy", saturate=False)
    assert maxrel(F / unit, tw["Fz_cpu_u"]) < 1e-13
    assert np.max(np.abs(P - tw["Pz_cpu_u"])) < 1e-14
    # independent implementation (Julia / OBE.jl)
    assert maxrel(F / unit, jl["Fz"]) < 1e-10
    Pjl = np.column_stack([jl[f"Pz_g{i + 1}"] for i in range(12)])
    assert np.max(np.abs(P - Pjl)) < 1e-10


def test_planted_mistakes_are_caught(mol, golden, hallas):
    """A wrong constant or a flipped polarisation in the oracle's inputs
    must fail the fixture comparison by orders of magnitude."""
    tw = golden("three_way_results.npz")
    unit = HBAR * mol.k * mol.Gamma / 2.0
    z = tw["z_arr"]
    # wrong hbar for the NumPy column: 1.6e-8 off, far above the 1e-13 bar
    F, _, _ = oracle.rate_eq_points(mol, hallas[0], np.zeros_like(z), z, B_HALLAS, hbar=HBAR_LEGACY,
                                    flavour="numpy")
    assert maxrel(F / unit, tw["Fz_cpu_s"]) > 1e-9
    # retro beam of component 0 with the wrong handedness
    bt = hallas[0]
    pols = bt.pols.copy()
    pols[1] = 2 - pols[1]
    bad = type(bt)(bt.dirs, bt.freqs, pols, bt.s0s, bt.comp, bt.split_sign, bt.n_comp)
    F, _, _ = oracle.rate_eq_points(mol, bad, np.zeros_like(z), z, B_HALLAS, hbar=HBAR, flavour="numpy")
    assert maxrel(F / unit, tw["Fz_cpu_s"]) > 1e-3
    # B gradient sign flip
    F, _, _ = oracle.rate_eq_points(mol, bt, np.zeros_like(z), z, -B_HALLAS, hbar=HBAR, flavour="numpy")
    assert maxrel(F / unit, tw["Fz_cpu_s"]) > 1e-3


def test_oracle_reproduces_reference_trajectories(mol, golden):
    tr = golden("trajectory_cpu_vs_gpu_results.npz")
    bt = beam_table(make_dc_mot_beams(mol, -0.20, 0.70, 1.0))
    dt = float(tr["dt"])
    n_steps = int(float(tr["t_max"]) / dt)
    zf, vf, na, _, _ = oracle.traj_euler(mol, bt, tr["ref_z0"], tr["ref_v0"], dt, n_steps, 0.16,
                                         hbar=HBAR_LEGACY)
    assert np.array_equal(zf, tr["ref_zf_jit"]) and np.array_equal(vf, tr["ref_vf_jit"])
    zf, vf, na, _, _ = oracle.traj_euler(mol, bt, tr["batch_z0"], tr["batch_v0"], dt, n_steps, 0.16,
                                         hbar=HBAR_LEGACY)
    assert np.array_equal(zf, tr["batch_zf_jit"]) and np.array_equal(vf, tr["batch_vf_jit"])
    assert np.array_equal(na, tr["batch_n_actual"])


def test_trajectory_escape_fill(mol):
    bt = beam_table(make_dc_mot_beams(mol, -0.20, 0.70, 1.0))
    zf, vf, na, zt, vt = oracle.traj_euler(mol, bt, [0.0], [-5.0], 1e-6, 5000, 0.16, hbar=HBAR_LEGACY,
                                           save_full=True)
    j = na[0]
    assert j < 5000
    assert abs(zt[0, j]) > 0.015 and abs(zt[0, j - 1]) <= 0.015
    assert np.all(zt[0, j:] == zt[0, j]) and np.all(vt[0, j:] == vt[0, j])
    assert zf[0] == zt[0, j] and vf[0] == vt[0, j]


def _lambda_molecule():
    """Two ground states, one excited state: |d|^2 = 0.3 (g0, sigma+),
    0.7 (g1, sigma-); g0 shifted by +2 MHz, Zeeman slopes +-0.1."""
    from molmot_amd.molecules.sroh import MolecularData
    tdm = np.zeros((3, 3, 3), dtype=complex)
    tdm[0, 2, 2] = np.sqrt(0.3)
    tdm[1, 2, 0] = np.sqrt(0.7)
    E = np.array([1.0e10 + 2e6, 1.0e10, 4.3e14])
    zz = np.diag([0.1, -0.1, 0.0]).astype(complex)
    Z0 = np.zeros((3, 3), dtype=complex)
    return MolecularData(energies=E, tdm=tdm, zeeman_x=Z0, zeeman_y=Z0, zeeman_z=zz,
                         n_ground=2, n_excited=1, n_states=3)


def test_lambda_system_closed_form(mol):
    """Lambda system: detailed balance p0 D0 b1 = p1 D1 b0 with D_i the
    saturated excitation rate out of g_i and b_i the branching ratios."""
    m = _lambda_molecule()
    beams = [LaserBeam(1.0, 4.3e14 - 1.0e10 - 2e6 - 3e6, 2, 2.0),
             LaserBeam(-1.0, 4.3e14 - 1.0e10 + 1e6, 0, 0.5)]
    bt = beam_table(beams)
    G, k = m.Gamma, m.k
    for v, z in [(0.0, 0.0), (1.3, 2e-3), (-4.0, -1e-3)]:
        F, p, R = oracle.rate_eq(m, bt, v, z, 0.2, hbar=HBAR, flavour="jit")
        B = 0.2 * z * 1e4
        Es = m.energies + B * m.zeeman_z_diag * G / (2 * np.pi)
        s = np.zeros(2)
        t = np.zeros(2)
        for b in beams:
            g = 0 if b.polarization == 2 else 1
            delta = 2 * np.pi * (b.freq_offset - k * b.kdir_z * v / (2 * np.pi) - (Es[2] - Es[g]))
            L = (G / 2) ** 2 / (delta ** 2 + (G / 2) ** 2)
            d2 = 0.3 if g == 0 else 0.7
            s[g] += b.s0 * d2 * L
            t[g] += b.kdir_z * b.s0 * d2 * L
        D = G / 2 * s / (1 + s)
        b0, b1 = 0.3, 0.7
        ratio = (D[1] * b0) / (D[0] * b1)      # p0 / p1
        p_exact = np.array([ratio, 1.0]) / (1.0 + ratio)
        F_exact = np.sum(p_exact * HBAR * k * G / 2 * t / (1 + s))
        R_exact = np.sum(p_exact * D)
        assert np.allclose(p, p_exact, rtol=1e-13, atol=0)
        assert abs(F - F_exact) <= 1e-12 * abs(F_exact)
        assert abs(R - R_exact) <= 1e-12 * R_exact


def test_steady_state_invariants_synthetic():
    """36+4 synthetic molecule: sum p = 1, p >= 0, and the NumPy and JIT
    flavours (different elimination codes) agree to rounding."""
    m = synthetic_molecule(36, 4, seed=3)
    bt = beam_table(make_dc_mot_beams(m, -0.5, 0.4, 2.0))
    rng = np.random.default_rng(0)
    v = rng.uniform(-10, 10, 64)
    z = rng.uniform(-5e-3, 5e-3, 64)
    Fj, Pj, Rj = oracle.rate_eq_points(m, bt, v, z, 0.2, hbar=HBAR, flavour="jit")
    Fn, Pn, Rn = oracle.rate_eq_points(m, bt, v, z, 0.2, hbar=HBAR, flavour="numpy")
    assert np.allclose(Pj.sum(1), 1.0, atol=1e-14) and np.all(Pj >= 0)
    assert maxrel(Fj, Fn) < 1e-11
    assert maxrel(Rj, Rn) < 1e-11
    assert np.max(np.abs(Pj - Pn)) < 1e-11


def test_dark_state_corner_cases(mol):
    """All s0 = 0: the JIT elimination skips every pivot and returns p = e0
    with F = R = 0; the NumPy path sees a singular matrix and returns the
    uniform distribution (molmot/obe/rate_equations.py:172-175)."""
    bt = beam_table(make_dc_mot_beams(mol, -0.2, 0.7, 0.0))
    F, p, R = oracle.rate_eq(mol, bt, 0.0, 1e-3, 0.16, flavour="jit")
    assert F == 0.0 and R == 0.0 and p[0] == 1.0 and np.all(p[1:] == 0.0)
    F, p, R = oracle.rate_eq(mol, bt, 0.0, 1e-3, 0.16, flavour="numpy")
    assert F == 0.0 and R == 0.0 and np.allclose(p, 1.0 / 12)
