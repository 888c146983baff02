"""Molecular tables from spectroscopic constants, and their CSV wire format.

Builds the 12 + 4 level tables of a linear 2Sigma+ molecule on the
X(N = 1) -> A 2Pi_1/2 (J' = 1/2) cycling transition:

1. ground basis: Hund's case (b), v = 0, S = I = 1/2, Lambda = 0, N = 1
   (12 states); H = B N^2 - D N^4 + gamma N.S + b_F I.S; eigenstates sorted
   by energy;
2. excited basis: Hund's case (a), Lambda = +-1, Sigma = -+1/2 (|P| = 1/2),
   J = 1/2 (8 states); H = A Lambda Sigma + B_A [J(J+1) - P^2] + (p + 2q)
   Lambda-doubling + T_A; the upper four eigenstates (the + parity doublet)
   are kept;
3. the excited eigenstates re-expressed in a case (b) basis (Lambda = -1..1,
   N = 0, 1, J = 1/2, N >= |K|: 16 states) through <b|a>, and the dipole
   table d[g, e, q] = <g| T^1_q(d) |e> carried into the eigenbases;
4. Zeeman tables: T^1_p(S) over the ground eigenstates,
   z = T_0, x = s (T_-1 - T_+1)/sqrt(2), y = s i (T_-1 + T_+1)/sqrt(2),
   with s = +1 (SrOH, CaF) or -1 (CaOH); the excited block is zero.

This is the pipeline of molmot/molecules/sroh.py:177-483 (SrOH),
caoh.py:89-371 (CaOH) and caf.py:90-333 (CaF), one function over a constants
record here.  ``load_caoh_from_sroh_structure`` / ``load_caf_from_sroh_structure``
(caoh.py:374-636, caf.py:336-576) keep steps 1, 2 and 4 but take the dipole
table from the SrOH CSV files (its angular structure is that of any
2Sigma+ -> 2Pi_1/2 transition), falling back to step 3 when they are absent.

``save_tables_csv`` / ``load_tables_csv`` write and read the wire format of
molmot/molecules/sroh.py:486-592 (energies; the dipole table as 3 n rows of
n columns, row = 3 state + q, real and imaginary parts in separate files;
each Zeeman matrix as n x n real and imaginary files) for any n_ground +
n_excited, so tables built here feed the solver kernels like the shipped
SrOH files do.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from ..constants import amu, c, gS, h, mu_B, wavenumber
from ..states import case_a, case_b
from ..states.basis import (HundsCaseA_LinearMolecule, HundsCaseB_LinearMolecule,
                            enumerate_states)
from ..states.hamiltonian import (Hamiltonian, compute_tdms, operator_to_matrix,
                                  overlap_caseb_casea)
from .sroh import SROH_DATA_DIR, MolecularData

__all__ = ["LinearMoleculeConstants", "SROH", "CAOH", "CAF", "build_tables",
           "build_sroh_hamiltonian", "build_caoh_hamiltonian", "build_caf_hamiltonian",
           "load_caoh_from_sroh_structure", "load_caf_from_sroh_structure",
           "save_tables_csv", "load_tables_csv", "zeeman_per_gauss", "CM_TO_HZ"]

CM_TO_HZ = c * 100.0


@dataclass(frozen=True)
class LinearMoleculeConstants:
    """Spectroscopic constants; every energy in Hz."""
    name: str
    B: float            # ground rotational constant
    D: float            # ground centrifugal distortion
    gamma: float        # spin-rotation
    bF: float           # Fermi contact
    T_A: float          # excited term energy
    B_A: float          # excited rotational constant
    A_so: float         # spin-orbit
    p2q: float          # Lambda-doubling p + 2q
    wavelength: float   # m
    gamma_hz: float     # natural linewidth / 2 pi
    mass_amu: float
    zeeman_sign: int    # sign of the x / y Zeeman convention

    @property
    def Gamma(self) -> float:
        return 2.0 * math.pi * self.gamma_hz

    @property
    def k(self) -> float:
        return wavenumber(self.wavelength)

    @property
    def mass(self) -> float:
        return self.mass_amu * amu


# molmot/molecules/sroh.py:56-81 (ground constants in cm^-1 except b_F)
SROH = LinearMoleculeConstants(
    "SrOH", B=0.24920 * CM_TO_HZ, D=0.0 * CM_TO_HZ, gamma=0.00242748 * CM_TO_HZ, bF=1.713e6,
    T_A=14542.573 * CM_TO_HZ, B_A=0.25389 * CM_TO_HZ, A_so=263.51741 * CM_TO_HZ,
    p2q=(-0.14320 + 2 * -0.00020) * CM_TO_HZ, wavelength=687.0e-9, gamma_hz=6.4e6,
    mass_amu=105.0, zeeman_sign=1)

# molmot/molecules/caoh.py:54-86 (ground constants in MHz, scaled as there)
CAOH = LinearMoleculeConstants(
    "CaOH", B=10023.0841 * 1e6, D=1.154e-2 * 1e6, gamma=34.7593 * 1e6, bF=2.602 * 1e6,
    T_A=15998.122 * CM_TO_HZ, B_A=0.3412200 * CM_TO_HZ, A_so=66.8181 * CM_TO_HZ,
    p2q=(-0.04287 + 2 * -0.3257e-3) * CM_TO_HZ, wavelength=626.0e-9, gamma_hz=6.4e6,
    mass_amu=57.0, zeeman_sign=-1)

# molmot/molecules/caf.py:54-87
CAF = LinearMoleculeConstants(
    "CaF", B=10303.988 * 1e6, D=0.014060 * 1e6, gamma=39.65891 * 1e6, bF=122.5569 * 1e6,
    T_A=16526.750 * CM_TO_HZ, B_A=0.348781 * CM_TO_HZ, A_so=71.429 * CM_TO_HZ,
    p2q=(-0.044517 + 2 * -2.916e-4) * CM_TO_HZ, wavelength=606.0e-9, gamma_hz=8.3e6,
    mass_amu=59.0, zeeman_sign=1)


def _ground(k: LinearMoleculeConstants):
    basis = enumerate_states(HundsCaseB_LinearMolecule,
                             {"v1": 0, "v2": 0, "v3": 0, "S": 0.5, "I": 0.5,
                              "Lambda": 0, "ell": 0, "N": [1]})
    H = Hamiltonian(basis)
    H.add_operator("B", k.B, case_b.Rotation)
    H.add_operator("D", k.D, case_b.RotationDistortion)
    H.add_operator("gamma", k.gamma, case_b.SpinRotation)
    H.add_operator("bF", k.bF, case_b.Hyperfine_IS)
    H.evaluate()
    return basis, H, H.solve()


def _excited(k: LinearMoleculeConstants):
    basis = [s for s in enumerate_states(HundsCaseA_LinearMolecule,
                                         {"v1": 0, "v2": 0, "v3": 0, "ell": 0,
                                          "Lambda": [-1, 1], "I": 0.5, "S": 0.5, "J": [0.5]})
             if abs(s.P) == 0.5]
    H = Hamiltonian(basis)
    H.add_operator("Aso", k.A_so, case_a.SpinOrbit)
    H.add_operator("Be", k.B_A, case_a.Rotation)
    H.add_operator("p+2q", k.p2q, case_a.LambdaDoubling_p2q)
    H.evaluate()
    H.matrix[np.diag_indices(len(basis))] += k.T_A
    states = sorted(H.solve(), key=lambda s: s.E)
    return basis, H, states[4:]


def _excited_caseb_basis():
    return [s for s in enumerate_states(HundsCaseB_LinearMolecule,
                                        {"v1": 0, "v2": 0, "v3": 0, "S": 0.5, "I": 0.5,
                                         "Lambda": [-1, 0, 1], "ell": 0, "N": [0, 1]})
            if abs(s.J - 0.5) < 0.01 and s.N >= abs(s.K)]


def _tdm_from_scratch(g_states, g_basis, e_states, e_basis_a):
    e_basis_b = _excited_caseb_basis()
    ov = np.array([[overlap_caseb_casea(sb, sa) for sa in e_basis_a] for sb in e_basis_b])
    Ce_b = np.array([s.coeffs for s in e_states]) @ ov.conj().T
    Cg = np.array([s.coeffs for s in g_states])
    return compute_tdms(g_states, e_states, g_basis, e_basis_b, case_b.TDM,
                        ground_coeffs=Cg, excited_coeffs=Ce_b)


def _sroh_csv_tdm(data_dir: str) -> Optional[np.ndarray]:
    re_p = os.path.join(data_dir, "sroh_tdm_re.csv")
    if not os.path.exists(re_p):
        return None
    flat = np.loadtxt(re_p) + 1j * np.loadtxt(os.path.join(data_dir, "sroh_tdm_im.csv"))
    return np.ascontiguousarray(flat.reshape(16, 3, 16).transpose(0, 2, 1))


def build_tables(k: LinearMoleculeConstants, tdm_source: str = "scratch",
                 sroh_data_dir: Optional[str] = None, verbose: bool = False) -> MolecularData:
    """The 12 + 4 tables of molecule ``k``; ``tdm_source`` "scratch" (step 3)
    or "sroh_csv" (the SrOH dipole table, step 3 if its files are absent)."""
    g_basis, _, g_states = _ground(k)
    e_basis_a, _, e_states = _excited(k)
    ng, ne = len(g_states), len(e_states)
    n = ng + ne
    d = None
    if tdm_source == "sroh_csv":
        d = _sroh_csv_tdm(os.path.abspath(sroh_data_dir or SROH_DATA_DIR))
        if d is None and verbose:
            print("SrOH dipole CSV files not found; computing the dipole table from scratch")
    elif tdm_source != "scratch":
        raise ValueError(f"tdm_source must be 'scratch' or 'sroh_csv', not {tdm_source!r}")
    if d is None:
        d = _tdm_from_scratch(g_states, g_basis, e_states, e_basis_a)
    Cg = np.array([s.coeffs for s in g_states])
    T = {p: Cg.conj() @ operator_to_matrix(g_basis, case_b.Zeeman, p) @ Cg.T for p in (-1, 0, 1)}
    Zx = np.zeros((n, n), complex)
    Zy = np.zeros((n, n), complex)
    Zz = np.zeros((n, n), complex)
    s = float(k.zeeman_sign)
    Zz[:ng, :ng] = T[0]
    Zx[:ng, :ng] = s * (T[-1] - T[1]) / math.sqrt(2)
    Zy[:ng, :ng] = s * 1j * (T[-1] + T[1]) / math.sqrt(2)
    E = np.array([st.E for st in g_states] + [st.E for st in e_states])
    mol = MolecularData(energies=E, tdm=d, zeeman_x=Zx, zeeman_y=Zy, zeeman_z=Zz,
                        Gamma=k.Gamma, k=k.k, mass=k.mass, wavelength=k.wavelength,
                        n_ground=ng, n_excited=ne, n_states=n)
    if verbose:
        print(f"{k.name}: {ng}+{ne} states, transition "
              f"{(E[ng:].mean() - E[:ng].mean()) / 1e12:.6f} THz, "
              f"spin-rotation splitting {(E[4:ng].mean() - E[:4].mean()) / 1e6:.1f} MHz")
    return mol


def build_sroh_hamiltonian(verbose: bool = False) -> MolecularData:
    return build_tables(SROH, "scratch", verbose=verbose)


def build_caoh_hamiltonian(verbose: bool = False) -> MolecularData:
    return build_tables(CAOH, "scratch", verbose=verbose)


def build_caf_hamiltonian(verbose: bool = False) -> MolecularData:
    return build_tables(CAF, "scratch", verbose=verbose)


def load_caoh_from_sroh_structure(sroh_data_dir: Optional[str] = None,
                                  verbose: bool = False) -> MolecularData:
    return build_tables(CAOH, "sroh_csv", sroh_data_dir, verbose)


def load_caf_from_sroh_structure(sroh_data_dir: Optional[str] = None,
                                 verbose: bool = False) -> MolecularData:
    return build_tables(CAF, "sroh_csv", sroh_data_dir, verbose)


def zeeman_per_gauss(mol: MolecularData) -> MolecularData:
    """A copy of ``mol`` whose Zeeman tables carry the scaling of the CSV
    tables, 2 pi gS mu_B / (Gamma h 1e4), so that E(Hz) = B(G) * zeeman *
    Gamma / (2 pi) (molmot/molecules/sroh.py:95-99).  The builders above
    return the bare electron-spin matrices T^1(S), as the reference's do."""
    f = 2.0 * math.pi * gS * mu_B / (mol.Gamma * h * 1e4)
    return MolecularData(energies=np.array(mol.energies), tdm=np.array(mol.tdm),
                         zeeman_x=f * mol.zeeman_x, zeeman_y=f * mol.zeeman_y,
                         zeeman_z=f * mol.zeeman_z, Gamma=mol.Gamma, k=mol.k, mass=mol.mass,
                         wavelength=mol.wavelength, n_ground=mol.n_ground,
                         n_excited=mol.n_excited, n_states=mol.n_states)


def save_tables_csv(mol: MolecularData, out_dir: str, prefix: str) -> None:
    """Write ``mol`` in the CSV wire format as ``<prefix>_energies.csv`` etc."""
    os.makedirs(out_dir, exist_ok=True)
    n = mol.n_states
    p = lambda name: os.path.join(out_dir, f"{prefix}_{name}.csv")  # noqa: E731
    fmt = "%.17g"
    np.savetxt(p("energies"), np.asarray(mol.energies, float), fmt=fmt)
    flat = np.asarray(mol.tdm).transpose(0, 2, 1).reshape(3 * n, n)
    np.savetxt(p("tdm_re"), flat.real, fmt=fmt)
    np.savetxt(p("tdm_im"), flat.imag, fmt=fmt)
    for ax in "xyz":
        z = np.asarray(getattr(mol, f"zeeman_{ax}"))
        np.savetxt(p(f"zeeman_{ax}_re"), z.real, fmt=fmt)
        np.savetxt(p(f"zeeman_{ax}_im"), z.imag, fmt=fmt)


def load_tables_csv(data_dir: str, prefix: str, n_ground: int, n_excited: int,
                    Gamma: float, k: float, mass: float, wavelength: float) -> MolecularData:
    """Read tables written in the CSV wire format (any n_ground + n_excited)."""
    n = n_ground + n_excited
    p = lambda name: os.path.join(data_dir, f"{prefix}_{name}.csv")  # noqa: E731
    E = np.atleast_1d(np.loadtxt(p("energies")))
    if E.shape != (n,):
        raise ValueError(f"expected {n} energies, got {E.shape}")
    flat = np.loadtxt(p("tdm_re"), ndmin=2) + 1j * np.loadtxt(p("tdm_im"), ndmin=2)
    if flat.shape != (3 * n, n):
        raise ValueError(f"dipole table shape {flat.shape}, expected {(3 * n, n)}")
    d = np.ascontiguousarray(flat.reshape(n, 3, n).transpose(0, 2, 1))
    Z = {}
    for ax in "xyz":
        z = np.loadtxt(p(f"zeeman_{ax}_re"), ndmin=2) + 1j * np.loadtxt(p(f"zeeman_{ax}_im"), ndmin=2)
        if z.shape != (n, n):
            raise ValueError(f"Zeeman {ax} shape {z.shape}")
        Z[ax] = z
    return MolecularData(energies=E, tdm=d, zeeman_x=Z["x"], zeeman_y=Z["y"], zeeman_z=Z["z"],
                         Gamma=Gamma, k=k, mass=mass, wavelength=wavelength,
                         n_ground=n_ground, n_excited=n_excited, n_states=n)
