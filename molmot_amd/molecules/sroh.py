"""SrOH molecular tables: container, CSV loader and a seeded synthetic builder.

``MolecularData`` keeps the attribute names and derived fields of the
reference container (molmot/molecules/sroh.py:87-174) so existing callers can
pass either object; ``load_sroh_from_julia`` reads the same CSV wire format
(molmot/molecules/sroh.py:486-592: 16 energies in Hz, TDM as 48x16 re/im rows
with row = 3*state + q, Zeeman x/y/z as 16x16 re/im).  The SrOH tables
themselves ship with this package under ``data/sroh``.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from ..constants import amu, wavenumber

# SrOH transition constants, molmot/molecules/sroh.py:75-81
LAMBDA_M = 687.0e-9
GAMMA_HZ = 6.4e6
GAMMA_RAD = 2.0 * math.pi * GAMMA_HZ
MASS_AMU = 105.0
MASS_KG = MASS_AMU * amu
K_WAVE = wavenumber(LAMBDA_M)

SROH_DATA_DIR = os.path.join(os.path.dirname(os.path.dirname(__file__)), "data", "sroh")


@dataclass
class MolecularData:
    """Tables for one molecule (reference: molmot/molecules/sroh.py:87-174).

    energies (n_states,) Hz; tdm (n_states, n_states, 3) complex with
    q index 0,1,2 = sigma-, pi, sigma+; zeeman_{x,y,z} (n_states, n_states)
    scaled so that E_zeeman(Hz) = B(G) * zeeman * Gamma/(2 pi).
    Derived: d_squared (n_g, n_e, 3) = |tdm[g, n_g+e, q]|^2, zeeman_z_diag,
    and the mean transition frequencies omega_J12 / omega_J32 / omega_mean.
    """
    energies: np.ndarray
    tdm: np.ndarray
    zeeman_x: np.ndarray
    zeeman_y: np.ndarray
    zeeman_z: np.ndarray
    Gamma: float = GAMMA_RAD
    k: float = K_WAVE
    mass: float = MASS_KG
    wavelength: float = LAMBDA_M
    n_ground: int = 12
    n_excited: int = 4
    n_states: int = 16
    d_squared: Optional[np.ndarray] = field(default=None, repr=False)
    zeeman_z_diag: Optional[np.ndarray] = field(default=None, repr=False)
    omega_J12: float = 0.0
    omega_J32: float = 0.0
    omega_mean: float = 0.0

    def __post_init__(self):
        ng, ne = self.n_ground, self.n_excited
        if self.d_squared is None:
            block = self.tdm[:ng, ng:ng + ne, :]
            self.d_squared = np.ascontiguousarray(np.abs(block) ** 2, dtype=np.float64)
        if self.zeeman_z_diag is None:
            self.zeeman_z_diag = np.ascontiguousarray(np.real(np.diag(self.zeeman_z)),
                                                      dtype=np.float64)
        E_g = self.energies[:ng]
        E_e_mean = np.mean(self.energies[ng:])
        # Manifold means: with 12 ground states the reference takes the first
        # 8 as J=3/2 and the last 4 as J=1/2 (sroh.py:159-174), even though
        # the CSV orders 4 low + 8 high states; kept as is so beam
        # frequencies built from these means agree with the reference.
        if ng == 12:
            mean_j32 = np.mean(E_g[:8])
            mean_j12 = np.mean(E_g[8:])
        else:
            mean_j32 = mean_j12 = np.mean(E_g)
        self.omega_J12 = E_e_mean - mean_j12
        self.omega_J32 = E_e_mean - mean_j32
        self.omega_mean = E_e_mean - np.mean(E_g)


def _read_table(path: str) -> np.ndarray:
    return np.loadtxt(path)


def load_sroh_from_julia(data_dir: str = SROH_DATA_DIR, verbose: bool = False) -> MolecularData:
    """Read the SrOH CSV tables (molmot/molecules/sroh.py:486-592)."""
    n_states, n_ground, n_excited = 16, 12, 4
    p = lambda name: os.path.join(data_dir, name)  # noqa: E731
    energies = _read_table(p("sroh_energies.csv"))
    if energies.shape != (n_states,):
        raise ValueError(f"expected {n_states} energies, got {energies.shape}")
    tdm_flat = _read_table(p("sroh_tdm_re.csv")) + 1j * _read_table(p("sroh_tdm_im.csv"))
    if tdm_flat.shape != (3 * n_states, n_states):
        raise ValueError(f"TDM table shape {tdm_flat.shape}")
    # row index = 3*state + q
    tdm = np.ascontiguousarray(tdm_flat.reshape(n_states, 3, n_states).transpose(0, 2, 1))

    def zeeman(axis):
        z = _read_table(p(f"sroh_zeeman_{axis}_re.csv")) + 1j * _read_table(p(f"sroh_zeeman_{axis}_im.csv"))
        if z.shape != (n_states, n_states):
            raise ValueError(f"Zeeman {axis} shape {z.shape}")
        return z

    mol = MolecularData(energies=energies, tdm=tdm, zeeman_x=zeeman("x"),
                        zeeman_y=zeeman("y"), zeeman_z=zeeman("z"),
                        Gamma=GAMMA_RAD, k=K_WAVE, mass=MASS_KG, wavelength=LAMBDA_M,
                        n_ground=n_ground, n_excited=n_excited, n_states=n_states)
    if verbose:
        print(f"SrOH: {n_ground}+{n_excited} states, lambda={LAMBDA_M*1e9:.1f} nm, "
              f"Gamma=2pi*{GAMMA_HZ/1e6:.1f} MHz")
    return mol


load_sroh = load_sroh_from_julia


def synthetic_molecule(n_ground: int = 12, n_excited: int = 4, seed: int = 0,
                       g_factor: float = 0.5) -> MolecularData:
    """Seeded synthetic n_ground + n_excited level structure.

    Ground states sit in two hyperfine/spin-rotation manifolds (lower third and
    upper two thirds, ~55 MHz apart, a few MHz internal spread) at ~1.5e10 Hz;
    excited states share ~4.32e14 Hz.  Each ground state carries a magnetic
    quantum number m in {-1,0,+1} and couples to excited state e with the
    polarisation q = m_e - m_g if |q| <= 1; complex dipoles are random and
    normalised so that sum over (g, q) of |d|^2 = 1 for every excited state.
    The ground Zeeman block is Hermitian with diagonal g*m*(mu_B/h)/(Gamma/2pi)
    per Gauss plus weak random off-diagonal mixing; the excited block is zero,
    as in the SrOH tables.
    """
    rng = np.random.default_rng(seed)
    ns = n_ground + n_excited
    n_low = n_ground // 3
    E = np.empty(ns)
    E[:n_low] = 1.4868e10 + rng.uniform(0.0, 2e5, n_low)
    E[n_low:n_ground] = 1.4977e10 + rng.uniform(0.0, 2e6, n_ground - n_low)
    E[n_ground:] = 4.320351078685233e14
    m_g = np.array([(i % 3) - 1 for i in range(n_ground)])
    m_e = np.array([(i % 3) - 1 for i in range(n_excited)])
    tdm = np.zeros((ns, ns, 3), dtype=complex)
    for ie in range(n_excited):
        col = []
        for ig in range(n_ground):
            q = m_e[ie] - m_g[ig]
            if abs(q) <= 1:
                d = rng.normal() + 1j * rng.normal()
                tdm[ig, n_ground + ie, q + 1] = d
                col.append((ig, q + 1))
        norm = math.sqrt(sum(abs(tdm[ig, n_ground + ie, qi]) ** 2 for ig, qi in col))
        for ig, qi in col:
            tdm[ig, n_ground + ie, qi] /= norm
            tdm[n_ground + ie, ig, qi] = np.conj(tdm[ig, n_ground + ie, qi])
    # mu_B/h = 1.3996 MHz/G; scaled units: Hz per Gauss / (Gamma/2pi)
    scale = g_factor * 1.39962449361e6 / GAMMA_HZ
    zz = np.zeros((ns, ns), dtype=complex)
    zz[:n_ground, :n_ground] = np.diag(scale * m_g)
    off = 0.02 * scale * (rng.normal(size=(n_ground, n_ground)) + 1j * rng.normal(size=(n_ground, n_ground)))
    off = np.triu(off, 1)
    zz[:n_ground, :n_ground] += off + off.conj().T
    zx = np.zeros((ns, ns), dtype=complex)
    zy = np.zeros((ns, ns), dtype=complex)
    for ig in range(n_ground - 1):
        zx[ig, ig + 1] = zx[ig + 1, ig] = 0.5 * scale
        zy[ig, ig + 1] = -0.5j * scale
        zy[ig + 1, ig] = 0.5j * scale
    return MolecularData(energies=E, tdm=tdm, zeeman_x=zx, zeeman_y=zy, zeeman_z=zz,
                         n_ground=n_ground, n_excited=n_excited, n_states=ns)
