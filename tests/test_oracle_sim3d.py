"""Pin the 3D rate-equation oracle (CPU): two-level saturated closed form,
squared Wigner d^1 polarisation weights, the reference's own 3D tests
(tests/test_all_functions.py:534-555), steady-state residuals, and the
leapfrog integrator on a constant force."""
import os

import numpy as np
import pytest

from oracle import sim3d as s3
from molmot_amd.molecules.sroh import MolecularData
from molmot_amd.obe.beam_pairs import make_beam_pairs_dc4


def two_level():
    tdm = np.zeros((2, 2, 3), dtype=complex)
    tdm[0, 1, 2] = tdm[1, 0, 2] = 1.0
    Z = np.zeros((2, 2), dtype=complex)
    return MolecularData(energies=np.array([1.0e9, 4.0e14]), tdm=tdm, zeeman_x=Z, zeeman_y=Z, zeeman_z=Z,
                         n_ground=1, n_excited=1, n_states=2)


@pytest.mark.parametrize("det_G,s0,vz", [(-1.0, 1.0, 0.0), (0.5, 3.0, 2.0), (-2.0, 0.4, -3.0)])
def test_two_level_closed_form(det_G, s0, vz):
    m = two_level()
    G = m.Gamma
    f = m.energies[1] - m.energies[0] + det_G * G / (2 * np.pi)
    beams = [dict(k_hat=np.array([0.0, 0.0, 1.0]), freq=f, q=2, s0=s0, w=None)]
    F, R, p, _ = s3.force(m, beams, [0.0, 0.0, 1e-3], [0.0, 0.0, vz], 0.2)
    delta = 2 * np.pi * (f - m.k * vz / (2 * np.pi) - (m.energies[1] - m.energies[0]))
    L = (G / 2) ** 2 / (delta ** 2 + (G / 2) ** 2)
    rate = G / 2 * s0 * L / (1 + s0 * L)
    assert abs(R - rate) < 1e-12 * rate
    assert np.allclose(F, [0.0, 0.0, s3.HBAR * m.k * rate], rtol=1e-12, atol=0)
    assert p[0] == 1.0


def test_polarisation_weights_closed_form():
    assert np.allclose(s3.pol_weights(2, np.array([0, 0, 1.0])), [0, 0, 1])
    assert np.allclose(s3.pol_weights(2, np.array([0, 0, -1.0])), [1, 0, 0])
    assert np.allclose(s3.pol_weights(2, np.array([1.0, 0, 0])), [0.25, 0.5, 0.25])
    assert np.allclose(s3.pol_weights(1, np.array([0, 1.0, 0])), [0.5, 0.0, 0.5])
    assert np.allclose(s3.pol_weights(0, np.zeros(3)), [1 / 3] * 3)


def test_reference_3d_tests(mol):
    beams = s3.six_beams(make_beam_pairs_dc4(mol, -0.30, 0.40, 1.0))
    F0, _, _, _ = s3.force(mol, beams, np.zeros(3), np.zeros(3), 0.16)
    assert np.linalg.norm(F0) < 1e-20
    F1, _, _, _ = s3.force(mol, beams, np.array([0.0, 0.0, 1e-3]), np.zeros(3), 0.16)
    assert F1[2] < 0


def test_steady_state_residual(mol):
    beams = s3.six_beams(make_beam_pairs_dc4(mol, -0.5, 0.6, 2.0))
    for r, v in [([1e-3, -2e-3, 5e-4], [0.3, -1.0, 2.0]), ([0, 0, 3e-3], [0, 0, -4.0])]:
        _, _, p, M = s3.force(mol, beams, r, v, 0.16)
        assert abs(p.sum() - 1) < 1e-14 and p.min() >= 0
        assert np.max(np.abs(M @ p)) < 1e-9 * np.max(np.abs(M))


def test_leapfrog_constant_force():
    m = two_level()
    a = np.array([1.0, -2.0, 0.5])
    t, r, v = s3.leapfrog(m, [], np.array([1e-4, 0, 0]), np.array([0.1, 0.2, -0.3]), 1e-3, 1e-5, 0.0,
                          r_escape=1.0, force_fn=lambda rr, vv: a * m.mass)
    tt = t[:, None]
    assert np.allclose(r, np.array([1e-4, 0, 0]) + np.array([0.1, 0.2, -0.3]) * tt + 0.5 * a * tt ** 2,
                       rtol=0, atol=1e-15)
    assert np.allclose(v[1:], np.array([0.1, 0.2, -0.3]) + a * tt[1:], rtol=0, atol=1e-13)


# ------------------------------------------------------------ spin basis

def test_spin_tables_detect_spin_algebra(mol):
    from molmot_amd.molecules import builders as bl, synthetic_molecule
    assert s3.spin_tables(mol) is not None
    assert s3.spin_tables(bl.zeeman_per_gauss(bl.load_caf_from_sroh_structure())) is not None
    assert s3.spin_tables(synthetic_molecule(12, 4, seed=3)) is None


def test_spin_basis_diagonalises_zeeman_block(mol):
    """U = R(B) W is unitary and U^H H_Z U = diag(|B| lam0), eigenvalues
    ascending, for random fields and the special directions +z, -z, and
    B = 0 (identity).  The tables close the spin algebra to ~1e-11
    relative, which bounds the residuals (5e-10 of |B|, against |B| s/2 =
    0.22 |B| for the eigenvalues)."""
    tabs = s3.spin_tables(mol)
    rng = np.random.default_rng(4)
    fields = [rng.normal(size=3) * 5 for _ in range(20)] + [np.array([0, 0, 3.0]), np.array([0, 0, -2.0]),
                                                           np.array([1e-9, 0, -2.0])]
    for B in fields:
        w, U = s3.spin_eig(tabs, B)
        H = s3.zeeman_block(mol, B)
        Bm = np.linalg.norm(B)
        assert np.all(np.diff(w) >= 0)
        assert np.max(np.abs(U.conj().T @ U - np.eye(12))) < 1e-10
        assert np.max(np.abs(U.conj().T @ H @ U - np.diag(w))) < 5e-10 * Bm
        assert np.allclose(np.sort(np.linalg.eigvalsh(H)), w, atol=5e-10 * Bm)
    w, U = s3.spin_eig(tabs, np.zeros(3))
    assert np.all(w == 0) and np.array_equal(U, np.eye(12))


def test_spin_basis_equals_reference_on_axis(mol):
    """On the +z axis U = W, the eigh basis of Zz, which is the reference's
    eigh basis of Bz Zz up to the reference eigensolver's own precision
    inside the near-degenerate eigenspaces of the SrOH tables: their six-fold
    eigenvalues are split at 1e-12, and LAPACK's eigh of 2.7 Zz differs from
    its eigh of Zz by rotations of up to 2e-3 inside those spaces.  Observed:
    force within 5e-6 of the reference's largest, R and p within 7e-6
    (tests/golden/ref_force3d_axis.npz); bars 2e-5 and 3e-5."""
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_force3d_axis.npz"))
    pairs = [dict(freq=float(f), q_fwd=int(a), q_bwd=int(b), s0=float(s))
             for f, a, b, s in zip(d["pairs_freq"], d["pairs_q_fwd"], d["pairs_q_bwd"], d["pairs_s0"])]
    beams = s3.six_beams(pairs, float(d["beam_radius"]))
    scale = np.max(np.abs(d["F"]))
    for i in range(len(d["r"])):
        F, R, p, _ = s3.force(mol, beams, d["r"][i], d["v"][i], float(d["B_gradient"]),
                              Gamma_eff_factor=float(d["gamma_eff_factor"]), basis="spin")
        assert np.max(np.abs(F - d["F"][i])) < 2e-5 * scale, i
        assert abs(R - d["R"][i]) < 3e-5 * d["R"][i], i
        assert np.max(np.abs(p - d["P"][i])) < 3e-5, i
