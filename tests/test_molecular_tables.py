"""Molecular-table pipeline (molmot_amd.states, molmot_amd.molecules.builders).

Pinned three ways:
* the Wigner symbols against sympy's exact ones;
* every operator matrix, Hamiltonian, eigenvector set and built table against
  the reference's own outputs (tests/golden/ref_tables.npz, written by
  tests/golden/make_reference_fixtures.py) -- observed agreement is <= 2e-16;
* closed forms the physics fixes: the N = 1 ground energies (spin-rotation
  plus Fermi contact, the F = 1 pair from its 2 x 2 block), the
  A 2Pi_1/2 J = 1/2 Lambda-doublet, the spin algebra of the Zeeman tables
  ([Sx, Sy] = i Sz, S^2 = 3/4, S_z = M / 4 on the F = 2 manifold), the dipole
  sum rules and the unitarity of the case (a) -> (b) transformation.
"""
import itertools
import os
from fractions import Fraction as Fr

import numpy as np
import pytest

from molmot_amd.molecules import builders as bl
from molmot_amd.molecules import load_sroh_from_julia, SROH_DATA_DIR
from molmot_amd.states import case_a as ca, case_b as cb
from molmot_amd.states.basis import (HundsCaseA_LinearMolecule as CaseA,
                                     HundsCaseB_LinearMolecule as CaseB, enumerate_states)
from molmot_amd.states.hamiltonian import operator_to_matrix, overlap_caseb_casea
from molmot_amd.states.wigner import wigner3j, wigner6j, wigner9j

HERE = os.path.dirname(os.path.abspath(__file__))
REF = np.load(os.path.join(HERE, "golden", "ref_tables.npz"))

BUILDERS = [("sroh_build", bl.build_sroh_hamiltonian), ("caoh_build", bl.build_caoh_hamiltonian),
            ("caf_build", bl.build_caf_hamiltonian), ("caoh_load", bl.load_caoh_from_sroh_structure),
            ("caf_load", bl.load_caf_from_sroh_structure)]
CONSTS = [("sroh", bl.SROH), ("caoh", bl.CAOH), ("caf", bl.CAF)]


def _qn(basis, names):
    return np.array([[getattr(s, n) for n in names] for s in basis])


def _halves(n):
    return [Fr(k, 2) for k in range(n)]


def test_wigner_vs_sympy():
    sw = pytest.importorskip("sympy.physics.wigner")
    for j1, j2, j3 in itertools.product(_halves(5), repeat=3):
        for m1 in np.arange(-j1, j1 + 1):
            for m2 in np.arange(-j2, j2 + 1):
                m3 = -m1 - m2
                if abs(m3) > j3:
                    continue
                try:
                    r = float(sw.wigner_3j(j1, j2, j3, m1, m2, m3))
                except ValueError:
                    r = 0.0
                assert abs(wigner3j(j1, j2, j3, m1, m2, m3) - r) < 1e-15
    rng = np.random.default_rng(5)
    hv = _halves(6)
    for _ in range(1500):
        args = [hv[i] for i in rng.integers(0, len(hv), 6)]
        try:
            r = float(sw.wigner_6j(*args))
        except ValueError:
            r = 0.0
        assert abs(wigner6j(*args) - r) < 1e-15
    n9 = 0
    for a, b, d, e in itertools.product(_halves(3), repeat=4):
        for c_ in np.arange(abs(a - b), a + b + 1):
            for f in np.arange(abs(d - e), d + e + 1):
                for g in np.arange(abs(a - d), min(a + d, 2) + 1):
                    for hh in np.arange(abs(b - e), min(b + e, 2) + 1):
                        i_ = abs(c_ - f)
                        r = float(sw.wigner_9j(a, b, c_, d, e, f, g, hh, i_, prec=None))
                        assert abs(wigner9j(a, b, c_, d, e, f, g, hh, i_) - r) < 1e-14
                        n9 += r != 0
    assert n9 > 30


def test_case_b_operators_vs_reference():
    big = enumerate_states(CaseB, {"v1": 0, "v2": 0, "v3": 0, "S": 0.5, "I": 0.5,
                                   "Lambda": 0, "ell": 0, "N": [0, 1, 2]})
    assert np.array_equal(_qn(big, ("N", "J", "F", "M")), REF["big_b_qn"])
    for name in ("Rotation", "RotationDistortion", "SpinRotation", "Hyperfine_IS"):
        assert np.max(np.abs(operator_to_matrix(big, getattr(cb, name)) - REF[f"big_b_{name}"])) < 1e-14
    for p in (-1, 0, 1):
        assert np.max(np.abs(operator_to_matrix(big, cb.Zeeman, p) - REF[f"big_b_Zeeman_p{p + 1}"])) < 1e-14
        assert np.max(np.abs(operator_to_matrix(big, cb.zeeman_nuclear, p)
                             - REF[f"big_b_zeeman_nuclear_p{p + 1}"])) < 1e-14
    exc = [s for s in enumerate_states(CaseB, {"v1": 0, "v2": 0, "v3": 0, "S": 0.5, "I": 0.5,
                                               "Lambda": [-1, 0, 1], "ell": 0, "N": [0, 1, 2]})
           if s.N >= abs(s.K)]
    assert np.array_equal(_qn(exc, ("Lambda", "K", "N", "J", "F", "M")), REF["big_e_qn"])
    for p in (-1, 0, 1):
        T = np.array([[cb.TDM(a, b, p) for b in exc] for a in big])
        assert np.max(np.abs(T - REF[f"big_tdm_p{p + 1}"])) < 1e-14


def test_case_a_operators_and_overlap_vs_reference():
    big_a = [s for s in enumerate_states(CaseA, {"v1": 0, "v2": 0, "v3": 0, "ell": 0, "Lambda": [-1, 1],
                                                 "I": 0.5, "S": 0.5, "J": [0.5, 1.5]})
             if abs(s.P) <= s.J]
    assert np.array_equal(_qn(big_a, ("Lambda", "Sigma", "J", "P", "F", "M")), REF["big_a_qn"])
    for name in ("Rotation", "SpinOrbit", "LambdaDoubling_p2q"):
        M = np.array([[getattr(ca, name)(a, b) for b in big_a] for a in big_a])
        assert np.max(np.abs(M - REF[f"big_a_{name}"])) < 1e-14
    exc = [s for s in enumerate_states(CaseB, {"v1": 0, "v2": 0, "v3": 0, "S": 0.5, "I": 0.5,
                                               "Lambda": [-1, 0, 1], "ell": 0, "N": [0, 1, 2]})
           if s.N >= abs(s.K)]
    O = np.array([[overlap_caseb_casea(b, a) for a in big_a] for b in exc])
    assert np.max(np.abs(O - REF["big_overlap_ba"])) < 1e-14
    # case (a) states with |Lambda| = 1 are spanned by the case (b) states of
    # the same Lambda, J, F, M: the transformation is unitary on them
    assert np.max(np.abs(O.conj().T @ O - np.eye(len(big_a)))) < 1e-13


def test_enumeration_requires_unbounded_numbers():
    with pytest.raises(ValueError):
        enumerate_states(CaseB, {"S": 0.5, "I": 0.5})
    with pytest.raises(ValueError):
        enumerate_states(CaseA, {"Lambda": [-1, 1], "S": 0.5, "I": 0.5})


@pytest.mark.parametrize("tag,k", CONSTS)
def test_builder_intermediates_vs_reference(tag, k):
    g_basis, Hg, g_states = bl._ground(k)
    e_basis, He, e_states = bl._excited(k)
    assert np.array_equal(_qn(g_basis, ("N", "J", "F", "M")), REF["ground_basis_qn"])
    assert np.array_equal(_qn(e_basis, ("Lambda", "Sigma", "J", "P", "F", "M")), REF["excited_a_qn"])
    assert np.max(np.abs(Hg.matrix - REF[f"{tag}_Hg"])) <= 1e-15 * np.max(np.abs(REF[f"{tag}_Hg"]))
    assert np.max(np.abs(He.matrix - REF[f"{tag}_He"])) <= 1e-15 * np.max(np.abs(REF[f"{tag}_He"]))
    assert np.max(np.abs(np.array([s.coeffs for s in g_states]) - REF[f"{tag}_Cg"])) < 1e-12
    assert np.max(np.abs(np.array([s.coeffs for s in e_states]) - REF[f"{tag}_Ce"])) < 1e-12
    Eall = np.sort(np.linalg.eigvalsh(He.matrix))
    assert np.max(np.abs(Eall - REF[f"{tag}_Ee_all"])) <= 1e-15 * Eall.max()
    for p in (-1, 0, 1):
        assert np.max(np.abs(operator_to_matrix(g_basis, cb.Zeeman, p) - REF[f"ground_zeeman_p{p + 1}"])) < 1e-14
        assert np.max(np.abs(operator_to_matrix(g_basis, cb.zeeman_nuclear, p)
                             - REF[f"ground_zeeman_nuclear_p{p + 1}"])) < 1e-14
        eb = bl._excited_caseb_basis()
        T = np.array([[cb.TDM(a, b, p) for b in eb] for a in g_basis])
        assert np.max(np.abs(T - REF[f"tdm_basis_p{p + 1}"])) < 1e-14
    eb = bl._excited_caseb_basis()
    assert np.array_equal(_qn(eb, ("Lambda", "K", "N", "J", "F", "M")), REF["excited_b_qn"])
    O = np.array([[overlap_caseb_casea(b, a) for a in e_basis] for b in eb])
    assert np.max(np.abs(O - REF["overlap_ba"])) < 1e-15


@pytest.mark.parametrize("tag,fn", BUILDERS)
def test_built_tables_vs_reference(tag, fn):
    m = fn()
    assert (m.n_ground, m.n_excited, m.n_states) == (12, 4, 16)
    E = REF[f"{tag}_energies"]
    assert np.max(np.abs(m.energies - E)) <= 1e-15 * np.max(np.abs(E))
    for key in ("tdm", "zeeman_x", "zeeman_y", "zeeman_z"):
        assert np.max(np.abs(np.asarray(getattr(m, key)) - REF[f"{tag}_{key}"])) < 1e-12, key
    assert np.allclose([m.Gamma, m.k, m.mass, m.wavelength], REF[f"{tag}_scalars"], rtol=1e-15, atol=0)


@pytest.mark.parametrize("tag,k", CONSTS)
def test_energies_closed_form(tag, k):
    """N = 1, S = I = 1/2: E(F=2) = 2B - 4D + gamma/2 + bF/4,
    E(F=0) = 2B - 4D - gamma + bF/4, and the F = 1 pair the eigenvalues of
    [[-gamma - bF/12, sqrt(2) bF/3], [sqrt(2) bF/3, gamma/2 - 5 bF/12]] + 2B - 4D
    (projection theorem on <I.S>); A 2Pi_1/2 J = 1/2: T - A/2 + B_A/2 -+ |p + 2q|/2,
    four states each (F = 0, 1)."""
    m = bl.build_tables(k)
    base = 2 * k.B - 4 * k.D
    x = np.sqrt(2.0) * k.bF / 3
    pair = np.linalg.eigvalsh(np.array([[-k.gamma - k.bF / 12, x], [x, k.gamma / 2 - 5 * k.bF / 12]]))
    expect = np.sort(np.r_[[base + k.gamma / 2 + k.bF / 4] * 5, [base - k.gamma + k.bF / 4],
                          [base + pair[0]] * 3, [base + pair[1]] * 3])
    assert np.max(np.abs(np.sort(m.energies[:12]) - expect)) < 1e-6 * max(abs(k.bF), abs(k.gamma)) + 1e-5
    up = k.T_A - k.A_so / 2 + k.B_A / 2 + abs(k.p2q) / 2
    assert np.max(np.abs(m.energies[12:] - up)) <= 2e-15 * up
    # the dropped lower doublet lies |p + 2q| below
    e_basis, He, _ = bl._excited(k)
    lo = np.sort(np.linalg.eigvalsh(He.matrix))[:4]
    assert np.max(np.abs(lo - (up - abs(k.p2q)))) <= 2e-15 * up


@pytest.mark.parametrize("tag,fn", BUILDERS)
def test_zeeman_spin_algebra_and_dipole_sum_rules(tag, fn):
    m = fn()
    ng = m.n_ground
    X, Y, Z = (np.asarray(getattr(m, f"zeeman_{a}"))[:ng, :ng] for a in "xyz")
    for A in (X, Y, Z):
        assert np.max(np.abs(A - A.conj().T)) < 1e-15
    assert np.max(np.abs(X @ Y - Y @ X - 1j * Z)) < 1e-14
    assert np.max(np.abs(X @ X + Y @ Y + Z @ Z - 0.75 * np.eye(ng))) < 1e-14
    for A in (X, Y, Z):
        assert np.all(np.asarray(getattr(m, "zeeman_x"))[ng:, :] == 0)
    # F = 2 manifold (the five degenerate states of E(F=2)): S_z = M / 4
    E = m.energies[:ng]
    f2 = np.where(np.abs(E - np.max(E[[0, 4, 7, 11]])) < 1.0)[0]
    if len(f2) == 5:
        assert np.allclose(np.sort(np.linalg.eigvalsh(Z[np.ix_(f2, f2)])),
                           [-0.5, -0.25, 0.0, 0.25, 0.5], atol=1e-14)
    d = m.tdm
    assert np.allclose(m.d_squared.sum(axis=(0, 2)), 1.0, atol=1e-13)
    assert np.allclose(m.d_squared.sum(axis=(0, 1)), 4.0 / 3.0, atol=1e-13)
    if tag.endswith("build"):
        for q in range(3):
            assert np.array_equal(d[ng:, :ng, 2 - q], d[:ng, ng:, q].conj().T)
    else:   # the SrOH CSV table holds the ground -> excited block only
        assert np.all(d[ng:, :ng] == 0)
    assert np.all(d[:ng, :ng] == 0) and np.all(d[ng:, ng:] == 0)


def test_caoh_sign_convention():
    """CaOH's x / y tables carry the opposite sign to SrOH's and CaF's
    (molmot/molecules/caoh.py:330-331)."""
    a = bl.build_tables(bl.CAOH)
    b = bl.build_tables(bl.LinearMoleculeConstants(**{**bl.CAOH.__dict__, "zeeman_sign": 1}))
    assert np.array_equal(a.zeeman_x, -b.zeeman_x) and np.array_equal(a.zeeman_y, -b.zeeman_y)
    assert np.array_equal(a.zeeman_z, b.zeeman_z)


def test_sroh_built_vs_julia_tables():
    """The from-scratch SrOH ground structure is that of the shipped tables to
    0.4 MHz (those include the dipolar hyperfine term and a small bias field);
    the dipole sum rules agree."""
    m = bl.build_sroh_hamiltonian()
    j = load_sroh_from_julia()
    gm = np.sort(m.energies[:12] - m.energies[:12].mean())
    gj = np.sort(j.energies[:12] - j.energies[:12].mean())
    assert np.max(np.abs(gm - gj)) < 0.5e6
    # optical transition: the Julia package's term energy differs by 3.8 GHz (9e-6)
    assert abs((m.energies[12:].mean() - m.energies[:12].mean())
               / (j.energies[12:].mean() - j.energies[:12].mean()) - 1) < 2e-5
    assert np.allclose(m.d_squared.sum(axis=(0, 1)), j.d_squared.sum(axis=(0, 1)), atol=1e-12)


def test_zeeman_per_gauss_matches_csv_scale():
    """Scaled, the built SrOH Zeeman eigenvalues are those of the CSV tables
    (+-gS mu_B / (2 h Gamma/2pi) per gauss for the stretched states)."""
    m = bl.zeeman_per_gauss(bl.build_sroh_hamiltonian())
    j = load_sroh_from_julia()
    ev = lambda z: np.linalg.eigvalsh(np.asarray(z)[:12, :12])  # noqa: E731
    assert np.allclose(ev(m.zeeman_z)[[0, -1]], ev(j.zeeman_z)[[0, -1]], rtol=1e-6)


def test_csv_round_trip(tmp_path):
    m = bl.load_caf_from_sroh_structure()
    bl.save_tables_csv(m, str(tmp_path), "caf")
    r = bl.load_tables_csv(str(tmp_path), "caf", 12, 4, m.Gamma, m.k, m.mass, m.wavelength)
    for key in ("energies", "tdm", "zeeman_x", "zeeman_y", "zeeman_z", "d_squared", "zeeman_z_diag"):
        assert np.array_equal(np.asarray(getattr(r, key)), np.asarray(getattr(m, key))), key
    assert r.omega_J12 == m.omega_J12 and r.omega_J32 == m.omega_J32
    s = bl.load_tables_csv(SROH_DATA_DIR, "sroh", 12, 4, 0.0, 0.0, 0.0, 0.0)
    j = load_sroh_from_julia()
    for key in ("energies", "tdm", "zeeman_x", "zeeman_y", "zeeman_z"):
        assert np.array_equal(np.asarray(getattr(s, key)), np.asarray(getattr(j, key)))
    with pytest.raises(ValueError):
        bl.load_tables_csv(str(tmp_path), "caf", 12, 5, 1.0, 1.0, 1.0, 1.0)


def test_build_tables_rejects_unknown_source():
    with pytest.raises(ValueError):
        bl.build_tables(bl.CAF, tdm_source="julia")
