"""DC and RF MOT simulators with the reference's constructor and ``force``
signatures (molmot/mot/simulator.py:17-162), evaluated on MI355X.

Each simulator keeps a device plan (the beam table uploaded once) and adds
``force_batch(v, z)`` for many points in one kernel launch.  ``hbar`` selects
the constant of the force: CODATA (molmot/constants.py:18) by default,
``constants.HBAR_LEGACY`` reproduces the reference's Numba/CUDA kernels.
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np
import torch

from ..hip.plan import HIPMolData
from ..hip.rate_equations import rate_eq_points_device
from ..obe.fields import make_dc_mot_beams, make_rf_mot_beams


def _points(plan, v, z, B):
    dev = torch.device("cuda", torch.cuda.current_device())
    shape = np.broadcast_shapes(np.shape(v), np.shape(z))
    v = np.array(np.broadcast_to(np.asarray(v, dtype=np.float64), shape))
    z = np.array(np.broadcast_to(np.asarray(z, dtype=np.float64), shape))
    F, P, R = rate_eq_points_device(plan, torch.from_numpy(v.ravel()).to(dev),
                                    torch.from_numpy(z.ravel()).to(dev), B)
    shape = v.shape
    return (F.cpu().numpy().reshape(shape), P.cpu().numpy().reshape(shape + (plan.n_ground,)),
            R.cpu().numpy().reshape(shape))


class DCMOTSimulator:
    """Static-field DC MOT (molmot/mot/simulator.py:93-162)."""

    def __init__(self, mol_data, delta_Gamma: float = -0.20, split_Gamma: float = 0.70,
                 s0: float = 1.0, B_gradient: float = 0.16, Gamma_eff_factor: float = 1.0,
                 hbar: Optional[float] = None):
        self.mol_data = mol_data
        self.delta_Gamma = delta_Gamma
        self.split_Gamma = split_Gamma
        self.s0 = s0
        self.B_gradient = B_gradient
        self.Gamma_eff_factor = Gamma_eff_factor
        self.hbar = hbar
        self._beams = make_dc_mot_beams(mol_data, delta_Gamma=delta_Gamma, split_Gamma=split_Gamma, s0=s0)
        self._plan_key = None
        self._plan_obj = None

    def update_beams(self):
        """Rebuild beams after parameter changes (molmot/mot/simulator.py:156-162)."""
        self._beams = make_dc_mot_beams(self.mol_data, delta_Gamma=self.delta_Gamma,
                                        split_Gamma=self.split_Gamma, s0=self.s0)
        self._plan_obj = None

    @property
    def _hip_plan(self) -> HIPMolData:
        key = (id(self._beams), float(self.Gamma_eff_factor), self.hbar)
        if self._plan_obj is None or self._plan_key != key:
            self._plan_obj = HIPMolData(self.mol_data, self._beams, self.Gamma_eff_factor, hbar=self.hbar)
            self._plan_key = key
        return self._plan_obj

    def force_batch(self, v, z):
        """Broadcast (v, z) arrays -> F [N], pop (..., n_g), R [1/s]."""
        return _points(self._hip_plan, v, z, self.B_gradient)

    def force(self, v: float, z: float) -> Tuple[float, np.ndarray, float]:
        F, P, R = self.force_batch(np.array([v]), np.array([z]))
        return float(F[0]), P[0], float(R[0])


class RFMOTSimulator:
    """RF MOT: average of the two switching phases (molmot/mot/simulator.py:17-90)."""

    def __init__(self, mol_data, delta_Gamma: float = -1.0, s0: float = 1.0, B_gradient: float = 0.16,
                 Gamma_eff_factor: float = 1.0, hbar: Optional[float] = None):
        self.mol_data = mol_data
        self.delta_Gamma = delta_Gamma
        self.s0 = s0
        self.B_gradient = B_gradient
        self.Gamma_eff_factor = Gamma_eff_factor
        self.hbar = hbar
        self._plans = None
        self._plans_key = None

    def _phase_plans(self):
        key = (self.delta_Gamma, self.s0, float(self.Gamma_eff_factor), self.hbar)
        if self._plans is None or self._plans_key != key:
            b0 = make_rf_mot_beams(self.mol_data, delta_Gamma=self.delta_Gamma, s0=self.s0, phase=0)
            b1 = make_rf_mot_beams(self.mol_data, delta_Gamma=self.delta_Gamma, s0=self.s0, phase=1)
            self._plans = (HIPMolData(self.mol_data, b0, self.Gamma_eff_factor, hbar=self.hbar),
                           HIPMolData(self.mol_data, b1, self.Gamma_eff_factor, hbar=self.hbar))
            self._plans_key = key
        return self._plans

    def force_batch(self, v, z):
        p0, p1 = self._phase_plans()
        F0, P0, R0 = _points(p0, v, z, self.B_gradient)
        F1, P1, R1 = _points(p1, v, z, -self.B_gradient)
        return 0.5 * (F0 + F1), 0.5 * (P0 + P1), 0.5 * (R0 + R1)

    def force(self, v: float, z: float) -> Tuple[float, np.ndarray, float]:
        F, P, R = self.force_batch(np.array([v]), np.array([z]))
        return float(F[0]), P[0], float(R[0])
