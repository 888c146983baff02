"""Device-resident rate-equation plan (the counterpart of the reference's
CUDAMolData, molmot_cuda/rate_equations_cuda.py:399-441).

A plan uploads the molecular tables and the beam set once; every batched
call reuses it.  hbar is explicit: ``constants.hbar`` (CODATA, the NumPy
path) by default, ``constants.HBAR_LEGACY`` reproduces the reference's Numba
and CUDA paths.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ..constants import hbar as HBAR_CODATA
from ..obe.fields import BeamTable, beam_table
from . import _lib


def component_split_signs(bt: BeamTable) -> np.ndarray:
    """d freq / d split per frequency component, as the plan stores it.  The
    split gradient is formed per component (sum over components of the sign
    times the component's detuning gradient), so every beam of a component
    must carry the same split sign; a mixed component raises ValueError."""
    n_comp = max(bt.n_comp, 1)
    split = np.zeros(n_comp)
    seen = np.zeros(n_comp, dtype=bool)
    for b in range(bt.n_beams):
        c = int(bt.comp[b])
        if seen[c] and split[c] != bt.split_sign[b]:
            raise ValueError(f"frequency component {c} mixes split signs {split[c]} and {bt.split_sign[b]}")
        split[c] = bt.split_sign[b]
        seen[c] = True
    return split


# kinds of plan-specialised kernels (HIPMolData.specialize)
FORWARD = 1
GRADIENT = 2
# particle-steps of a launch from which the "auto" rule builds the
# plan-specialised kernels of a kind (about 35 ms of generic forward time and
# 350 ms of generic gradient time on MI355X)
SPECIALIZE_MIN_WORK = {FORWARD: 5e8, GRADIENT: 2e9}


class HIPMolData:
    """Plan for one (molecule, beams, Gamma_eff, hbar).

    Mirrors CUDAMolData(mol_data, beams, Gamma_eff_factor); attributes
    n_ground, n_excited, n_states, n_beams, Gamma, k, Gamma_eff, hbar_val, mass.
    """

    def __init__(self, mol_data, beams, Gamma_eff_factor: float = 1.0, hbar: float | None = None,
                 specialize="auto"):
        _lib.require_gpu()
        if specialize not in ("auto", True, False):
            raise ValueError("specialize must be 'auto', True or False")
        # "auto": the trajectory launchers specialise the plan (specialize())
        # on its first finals-only launch of at least SPECIALIZE_MIN_WORK
        # particle-steps, the forward-sensitivity launchers likewise for the
        # gradient kernels; True: the trajectory kernels now; False: never
        self.auto_specialize = (FORWARD | GRADIENT) if specialize == "auto" else 0
        bt = beams if isinstance(beams, BeamTable) else beam_table(beams)
        self.mol_data = mol_data
        self.beams = bt
        self.n_ground = int(mol_data.n_ground)
        self.n_excited = int(mol_data.n_excited)
        self.n_states = self.n_ground + self.n_excited
        self.n_beams = bt.n_beams
        self.Gamma = float(mol_data.Gamma)
        self.k = float(mol_data.k)
        self.mass = float(mol_data.mass)
        self.Gamma_eff = float(Gamma_eff_factor)
        self.hbar_val = float(HBAR_CODATA if hbar is None else hbar)
        self._handle = ctypes.c_void_p()
        L = _lib.lib()
        E = np.ascontiguousarray(mol_data.energies, dtype=np.float64)
        zd = np.ascontiguousarray(mol_data.zeeman_z_diag, dtype=np.float64)
        d2 = np.ascontiguousarray(mol_data.d_squared, dtype=np.float64)
        if E.shape != (self.n_states,) or zd.shape != (self.n_states,) or \
                d2.shape != (self.n_ground, self.n_excited, 3):
            raise ValueError("molecular table shapes do not match n_ground/n_excited")
        n_comp = max(bt.n_comp, 1)
        split = component_split_signs(bt)
        keep = [E, zd, d2, bt.dirs, bt.freqs, bt.pols, bt.s0s, bt.comp, split]
        dp = lambda a: a.ctypes.data_as(_lib._dp)  # noqa: E731
        ip = lambda a: a.ctypes.data_as(_lib._ip)  # noqa: E731
        rc = L.mm_rate_plan_create(ctypes.byref(self._handle), self.n_ground, self.n_excited,
                                   self.n_beams, dp(E), dp(zd), dp(d2), dp(bt.dirs), dp(bt.freqs),
                                   ip(bt.pols), dp(bt.s0s), ip(bt.comp), dp(split), n_comp,
                                   self.Gamma, self.k, self.Gamma_eff, self.hbar_val)
        del keep
        _lib.check(rc, "mm_rate_plan_create")
        self.n_terms = int(L.mm_rate_plan_n_terms(self._handle))
        if specialize is True:
            self.specialize(FORWARD)

    @property
    def handle(self):
        return self._handle

    def block_ok(self, block: int) -> bool:
        """Whether the map / point kernels can run with ``block`` threads."""
        return block in (64, 128, 256) and 2 * self.n_ground * 8 * block <= 49152

    def set_block(self, block: int) -> None:
        """Threads per block of the map / point kernels (0: default)."""
        _lib.check(_lib.lib().mm_rate_plan_set_block(self._handle, int(block)), "mm_rate_plan_set_block")

    @property
    def block(self) -> int:
        return int(_lib.lib().mm_rate_plan_get_block(self._handle))

    def maybe_specialize(self, particle_steps: float, kind: int = FORWARD) -> None:
        """The "auto" rule: specialise on a launch of at least
        SPECIALIZE_MIN_WORK[kind] particle-steps (the build, ~1 s for the
        trajectory kernels and a few seconds for the gradient kernels, is then
        repaid within a launch or two: the specialised kernels take 35-55%
        less time).  Decided once per plan and kind."""
        if self.auto_specialize & kind and particle_steps >= SPECIALIZE_MIN_WORK[kind] and \
                not self.specialized & kind:
            self.auto_specialize &= ~kind
            if self.specializable(kind):
                self.specialize(kind)

    def specializable(self, kind: int = FORWARD) -> bool:
        """Whether the plan's tables admit the plan-specialised kernels of
        this kind: a packed plan (unit beam directions, n_excited <= 4, every
        (ground, excited) pair empty or driven by two +z and two -z beams, no
        dark ground state); for GRADIENT also one beam of each of at most
        four frequency components per pair."""
        return int(_lib.lib().mm_rate_plan_spec_source(self._handle, int(kind), None, 0)) > 0

    def specialize(self, kinds: int = FORWARD) -> None:
        """Compile (hipRTC; modules cached per process) kernels with this
        plan's tables as constants of the code: FORWARD, the 1D trajectory
        kernels (mm_traj_euler_1d, mm_traj_euler_1d_ws with or without the
        saved trajectory, mm_traj_euler_1d_noalloc), the force map / points, the
        Monte-Carlo temperature sums, the per-axis 3D finals and the
        ensemble statistics (bit-identical results), GRADIENT, the force-gradient kernels (trajectory_tangent,
        trajectory_vjp's segments, force_and_jacobian), or both (3).
        ``kinds=0`` returns the plan to the generic kernels."""
        kinds = int(kinds)
        _lib.check(_lib.lib().mm_rate_plan_specialize(self._handle, kinds), "mm_rate_plan_specialize")

    @property
    def specialized(self) -> int:
        """Mask of the kinds of specialised kernels the plan carries."""
        return int(_lib.lib().mm_rate_plan_is_specialized(self._handle))

    def spec_source(self, kind: int = FORWARD) -> str:
        """The generated source of the specialised kernels (for inspection)."""
        L = _lib.lib()
        n = int(L.mm_rate_plan_spec_source(self._handle, int(kind), None, 0))
        if n <= 0:
            _lib.check(n, "mm_rate_plan_spec_source")
        buf = ctypes.create_string_buffer(n)
        L.mm_rate_plan_spec_source(self._handle, int(kind), buf, n)
        return buf.value.decode()

    # RF MOT: kernels for this plan as the first switching phase and
    # `other` as the second (mm_rate_plan_specialize_rf)

    def rf_specializable(self, other: "HIPMolData") -> bool:
        """Whether (self, other) admit the plan-specialised RF kernels (unit
        beam directions, equal state counts, no dark ground state)."""
        return int(_lib.lib().mm_rate_plan_spec_source_rf(self._handle, other.handle, None, 0)) > 0

    def specialize_rf(self, other: "HIPMolData | None") -> None:
        """Compile the RF kernels of mm_traj_euler_rf_1d and
        mm_traj_vsq_rf_1d with both phase plans' tables as constants
        (bit-identical results); ``None`` releases them."""
        h = other.handle if other is not None else None
        _lib.check(_lib.lib().mm_rate_plan_specialize_rf(self._handle, h), "mm_rate_plan_specialize_rf")

    def specialized_rf(self, other: "HIPMolData") -> bool:
        return bool(_lib.lib().mm_rate_plan_is_specialized_rf(self._handle, other.handle))

    def maybe_specialize_rf(self, other: "HIPMolData", particle_steps: float) -> None:
        """The "auto" rule of maybe_specialize for the pair (self, other):
        on a launch of at least SPECIALIZE_MIN_WORK[FORWARD] particle-steps,
        once per pair of plans."""
        tried = self.__dict__.setdefault("_rf_tried", set())
        if not self.auto_specialize & FORWARD or particle_steps < SPECIALIZE_MIN_WORK[FORWARD] or \
                id(other) in tried or self.specialized_rf(other):
            return
        tried.add(id(other))
        if self.rf_specializable(other):
            self.specialize_rf(other)

    def spec_source_rf(self, other: "HIPMolData") -> str:
        """The generated source of the RF kernels (for inspection)."""
        L = _lib.lib()
        n = int(L.mm_rate_plan_spec_source_rf(self._handle, other.handle, None, 0))
        if n <= 0:
            _lib.check(n, "mm_rate_plan_spec_source_rf")
        buf = ctypes.create_string_buffer(n)
        L.mm_rate_plan_spec_source_rf(self._handle, other.handle, buf, n)
        return buf.value.decode()

    def close(self):
        if getattr(self, "_handle", None) and self._handle.value:
            _lib.check(_lib.lib().mm_rate_plan_destroy(self._handle), "mm_rate_plan_destroy")
            self._handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# the reference's class name, for callers that construct it directly
CUDAMolData = HIPMolData
