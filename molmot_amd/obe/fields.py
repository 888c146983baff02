"""Laser beams of the 1D MOT models and the flat beam table the kernels use.

``LaserBeam``, ``make_dc_mot_beams`` and ``make_rf_mot_beams`` keep the
reference's signatures and conventions (molmot/obe/fields.py:22-58, 76-136,
139-202): polarisation index 0/1/2 = sigma-/pi/sigma+, a retro-reflected beam
swaps sigma+ and sigma-, and ``kdir_z`` is the k-vector's z projection.

``BeamTable`` is this package's own addition: the flat arrays the device plan
uploads, plus the parameter map used by the adjoint kernels.  Each beam
belongs to a frequency component ``comp`` and carries a split sign
``split_sign`` so that

    freq_b = base_comp + detuning_comp + split_sign_comp * split

which is how both the DC MOT (molmot/obe/fields.py:174-179: one detuning,
+-split by polarisation) and the four-detuning Hallas configuration
(validation/sroh_redmot_three_way.py:67-103) are parameterised.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from ..constants import h, c


@dataclass
class LaserBeam:
    """One beam / frequency component (molmot/obe/fields.py:22-58)."""
    direction: np.ndarray
    freq_offset: float
    polarization: int
    s0: float
    comp: int = -1            # frequency component index (for gradients)
    split_sign: float = 0.0   # d freq / d split, in units of the split

    @property
    def kdir_z(self) -> float:
        d = self.direction
        if isinstance(d, np.ndarray):
            return float(d[2])
        return float(d)

    def __post_init__(self):
        if isinstance(self.direction, (int, float)):
            self.direction = np.array([0.0, 0.0, float(self.direction)])
        self.direction = np.asarray(self.direction, dtype=float)


def _retro(q: int) -> int:
    return 2 - q if q != 1 else 1


def make_dc_mot_beams(mol_data, delta_Gamma: float = -0.20, split_Gamma: float = 0.70,
                      s0: float = 1.0, n_freq: int = 4) -> List[LaserBeam]:
    """Four-frequency DC MOT, 8 beams (molmot/obe/fields.py:139-202)."""
    G = mol_data.Gamma
    delta_hz = delta_Gamma * G / (2.0 * math.pi)
    split_hz = split_Gamma * G / (2.0 * math.pi)
    comps = [
        (mol_data.omega_J32 + delta_hz + split_hz, 2, +1.0),
        (mol_data.omega_J32 + delta_hz - split_hz, 0, -1.0),
        (mol_data.omega_J12 + delta_hz + split_hz, 2, +1.0),
        (mol_data.omega_J12 + delta_hz - split_hz, 0, -1.0),
    ]
    beams = []
    for ic, (freq, q_fwd, sgn) in enumerate(comps):
        for direction in (+1, -1):
            q = q_fwd if direction > 0 else _retro(q_fwd)
            beams.append(LaserBeam(np.array([0.0, 0.0, float(direction)]), freq, q, s0,
                                   comp=ic, split_sign=sgn))
    return beams


def make_rf_mot_beams(mol_data, delta_Gamma: float = -1.0, s0: float = 1.0,
                      phase: int = 0, n_sidebands: int = 2) -> List[LaserBeam]:
    """RF MOT beams for one switching phase (molmot/obe/fields.py:76-136)."""
    G = mol_data.Gamma
    delta_hz = delta_Gamma * G / (2.0 * math.pi)
    f12 = mol_data.omega_J12 + delta_hz
    f32 = mol_data.omega_J32 + delta_hz
    q32, q12 = (2, 0) if phase == 0 else (0, 2)
    beams = []
    for ic, (freq, q_fwd) in enumerate([(f32, q32), (f12, q12)]):
        for direction in (+1, -1):
            q = q_fwd if direction > 0 else (2 - q_fwd)
            beams.append(LaserBeam(np.array([0.0, 0.0, float(direction)]), freq, q, s0,
                                   comp=ic, split_sign=0.0))
    return beams


# Hallas SrOH red-MOT configuration (validation/sroh_redmot_three_way.py:56-103)
HALLAS_DETUNINGS_MHZ = (-10.3, -5.9, -9.2, -29.9)
HALLAS_POWER_RATIOS = (0.50, 0.08, 0.15, 0.27)
HALLAS_POLS = (2, 0, 2, 0)
HALLAS_SPLIT_SIGNS = (+1.0, -1.0, +1.0, -1.0)
HALLAS_B_GRAD_GCM = 21.4


def hallas_saturations(Gamma: float, total_power_mW: float = 100.0,
                       beam_radius_m: float = 10e-3, wavelength_m: float = 687e-9,
                       ratios: Sequence[float] = HALLAS_POWER_RATIOS,
                       per_beam_divisor: float = 1.0) -> np.ndarray:
    """s0 per frequency component: I = 2P/(pi w^2) over Isat = pi h c Gamma/(3 lambda^3)
    (validation/sroh_redmot_three_way.py:85-90).  ``per_beam_divisor`` = 6
    gives the "total power over 6 beams" reading."""
    isat = np.pi * h * c * Gamma / (3.0 * wavelength_m ** 3)
    r = np.array(ratios, dtype=float)
    r /= r.sum()
    powers_W = total_power_mW * 1e-3 * r / per_beam_divisor
    intensities = powers_W * (2.0 / (np.pi * beam_radius_m ** 2))
    return intensities / isat


def make_hallas_beams(mol_data, detunings_MHz: Sequence[float] = HALLAS_DETUNINGS_MHZ,
                      sats: Optional[Sequence[float]] = None,
                      beam_radius_m: float = 10e-3, split_Hz: float = 0.0):
    """Hallas four-frequency beams: 4 components x (+z, -z)
    (validation/sroh_redmot_three_way.py:67-103).  Returns (beams, freqs, sats)."""
    n_g, n_e = mol_data.n_ground, mol_data.n_excited
    E = mol_data.energies
    E_e_last = E[n_g + n_e - 1]
    det_hz = np.array(detunings_MHz, dtype=float) * 1e6
    freqs = [E_e_last - E[0] + det_hz[0], E_e_last - E[0] + det_hz[1],
             E_e_last - E[9] + det_hz[2], E_e_last - E[9] + det_hz[3]]
    if sats is None:
        sats = hallas_saturations(mol_data.Gamma, beam_radius_m=beam_radius_m)
    beams = []
    for ic, (f, q_fwd, s0) in enumerate(zip(freqs, HALLAS_POLS, sats)):
        sgn = HALLAS_SPLIT_SIGNS[ic]
        for direction in (+1, -1):
            q = q_fwd if direction > 0 else _retro(q_fwd)
            beams.append(LaserBeam(np.array([0.0, 0.0, float(direction)]),
                                   f + sgn * split_Hz, q, float(s0), comp=ic, split_sign=sgn))
    return beams, freqs, np.asarray(sats, dtype=float)


@dataclass
class BeamTable:
    """Flat beam arrays (what the kernels and the oracle consume)."""
    dirs: np.ndarray
    freqs: np.ndarray
    pols: np.ndarray
    s0s: np.ndarray
    comp: np.ndarray
    split_sign: np.ndarray
    n_comp: int = 0

    @property
    def n_beams(self) -> int:
        return int(self.dirs.shape[0])


def beam_table(beams: Sequence[LaserBeam]) -> BeamTable:
    """Pack a beam list; beams without a component index get one per
    distinct frequency, in order of appearance."""
    dirs = np.array([b.kdir_z for b in beams], dtype=np.float64)
    freqs = np.array([b.freq_offset for b in beams], dtype=np.float64)
    pols = np.array([int(b.polarization) for b in beams], dtype=np.int32)
    s0s = np.array([b.s0 for b in beams], dtype=np.float64)
    comp = []
    seen = {}
    for b in beams:
        cidx = getattr(b, "comp", -1)
        if cidx is None or cidx < 0:
            key = float(b.freq_offset)
            if key not in seen:
                seen[key] = len(seen)
            cidx = seen[key]
        comp.append(int(cidx))
    comp = np.array(comp, dtype=np.int32)
    sign = np.array([float(getattr(b, "split_sign", 0.0)) for b in beams], dtype=np.float64)
    n_comp = int(comp.max()) + 1 if len(comp) else 0
    if np.any((pols < 0) | (pols > 2)):
        raise ValueError("beam polarisation index must be 0, 1 or 2")
    return BeamTable(dirs, freqs, pols, s0s, comp, sign, n_comp)
