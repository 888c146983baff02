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


# particle-steps of a finals-only trajectory launch from which the "auto"
# rule compiles the plan-specialised kernels (about 40 ms of generic-kernel
# time on MI355X)
SPECIALIZE_MIN_WORK = 5e8


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
        # particle-steps; True: now; False: never
        self.auto_specialize = specialize == "auto"
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
            self.specialize()

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

    def maybe_specialize(self, particle_steps: float) -> None:
        """The "auto" rule: specialise on a launch of at least
        SPECIALIZE_MIN_WORK particle-steps (the ~0.5 s compile is then repaid
        within a few launches: the specialised kernels take ~35% less time)."""
        if self.auto_specialize and particle_steps >= SPECIALIZE_MIN_WORK and not self.specialized:
            self.auto_specialize = False  # decide once per plan
            if self.specializable():
                self.specialize()

    def specializable(self) -> bool:
        """Whether the plan's tables admit the plan-specialised trajectory
        kernels (a packed plan: unit beam directions, n_excited <= 4, every
        (ground, excited) pair empty or driven by two +z and two -z beams,
        no dark ground state)."""
        return int(_lib.lib().mm_rate_plan_spec_source(self._handle, None, 0)) > 0

    def specialize(self, enable: bool = True) -> None:
        """Compile (hipRTC, ~1 s, cached per process) the 1D trajectory
        kernels with this plan's tables as constants of the code; later
        trajectory launches on the plan (finals only: mm_traj_euler_1d,
        mm_traj_euler_1d_ws, mm_traj_euler_1d_noalloc) run them.  The
        trajectories are bit-identical to the generic kernels'.
        ``enable=False`` returns the plan to the generic kernels."""
        _lib.check(_lib.lib().mm_rate_plan_specialize(self._handle, 1 if enable else 0),
                   "mm_rate_plan_specialize")

    @property
    def specialized(self) -> bool:
        return bool(_lib.lib().mm_rate_plan_is_specialized(self._handle))

    def spec_source(self) -> str:
        """The generated source of the specialised kernels (for inspection)."""
        L = _lib.lib()
        n = int(L.mm_rate_plan_spec_source(self._handle, None, 0))
        if n <= 0:
            _lib.check(n, "mm_rate_plan_spec_source")
        buf = ctypes.create_string_buffer(n)
        L.mm_rate_plan_spec_source(self._handle, buf, n)
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
