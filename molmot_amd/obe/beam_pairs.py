"""Counter-propagating beam-pair descriptions {'freq' [Hz], 'q_fwd', 'q_bwd',
's0'} used by the Lindblad and MCWF solvers (molmot/obe/obe_subdoppler.py:289-337).
Polarisation index q: 0 = sigma-, 1 = pi, 2 = sigma+."""
from __future__ import annotations

import numpy as np


def _hz(mol, x_gamma):
    return x_gamma * mol.Gamma / (2 * np.pi)


def make_beam_pairs_rf(mol_data, delta_Gamma, s0, phase=0):
    """Two pairs (J=3/2 and J=1/2 transitions) for one phase of the RF MOT
    cycle (obe_subdoppler.py:289-305): phase 0 drives J=3/2 with sigma+ and
    J=1/2 with sigma- along +k, phase 1 swaps them."""
    d = _hz(mol_data, delta_Gamma)
    q32, q12 = (2, 0) if phase == 0 else (0, 2)
    return [{"freq": mol_data.omega_J32 + d, "q_fwd": q32, "q_bwd": 2 - q32, "s0": s0},
            {"freq": mol_data.omega_J12 + d, "q_fwd": q12, "q_bwd": 2 - q12, "s0": s0}]


def make_beam_pairs_dc4(mol_data, delta_Gamma, split_Gamma, s0):
    """Four-frequency DC MOT (obe_subdoppler.py:308-323): each of J=3/2 and
    J=1/2 gets a sigma+ component at delta + split and a sigma- component at
    delta - split."""
    d, sp = _hz(mol_data, delta_Gamma), _hz(mol_data, split_Gamma)
    out = []
    for base in (mol_data.omega_J32, mol_data.omega_J12):
        out.append({"freq": base + d + sp, "q_fwd": 2, "q_bwd": 0, "s0": s0})
        out.append({"freq": base + d - sp, "q_fwd": 0, "q_bwd": 2, "s0": s0})
    return out


def make_beam_pairs_blue(mol_data, delta_Gamma, delta2_Gamma, s0):
    """Blue-detuned Lambda configuration (obe_subdoppler.py:326-337)."""
    d, d2 = _hz(mol_data, delta_Gamma), _hz(mol_data, delta2_Gamma)
    return [{"freq": mol_data.omega_J32 + d, "q_fwd": 0, "q_bwd": 2, "s0": s0},
            {"freq": mol_data.omega_J12 + d + d2, "q_fwd": 2, "q_bwd": 0, "s0": s0}]
