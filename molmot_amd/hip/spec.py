"""Plan-specialised kernels without a device: the generated source of a
(molecule, beams) plan and its hipRTC compilation for gfx950, as the library
builds them at mm_rate_plan_specialize (see HIPMolData.specialize).

    src = spec_source(mol, beams)            # str, or None if not specialisable
    gsrc = spec_source(mol, beams, kind=2)   # the forward-sensitivity kernels
    code = compile_source(src)               # gfx950 code object (ELF bytes)
    counts = fp64_counts(code, "mms_traj_queue_k1")   # FP64 VALU instructions

No compute runs here: these are for inspection, for checking the generator
on a host without a GPU, and for the executed-operation count of the
specialised step (fp64_counts disassembles with ROCm's llvm-objdump).
"""
from __future__ import annotations

import ctypes
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np

from ..constants import hbar as HBAR_CODATA
from ..obe.fields import BeamTable, beam_table
from . import _lib


def spec_source(mol_data, beams, Gamma_eff_factor: float = 1.0, hbar: float | None = None, kind: int = 1):
    """Generated source of the specialised kernels of this kind (1: trajectory,
    2: forward sensitivities) for these tables, or None when the plan cannot
    have them (MM_EUNSUPPORTED)."""
    from .plan import component_split_signs
    bt = beams if isinstance(beams, BeamTable) else beam_table(beams)
    E = np.ascontiguousarray(mol_data.energies, dtype=np.float64)
    zd = np.ascontiguousarray(mol_data.zeeman_z_diag, dtype=np.float64)
    d2 = np.ascontiguousarray(mol_data.d_squared, dtype=np.float64)
    split = component_split_signs(bt)
    dp = lambda a: a.ctypes.data_as(_lib._dp)  # noqa: E731
    ip = lambda a: a.ctypes.data_as(_lib._ip)  # noqa: E731
    args = (int(mol_data.n_ground), int(mol_data.n_excited), bt.n_beams, dp(E), dp(zd), dp(d2), dp(bt.dirs),
            dp(bt.freqs), ip(bt.pols), dp(bt.s0s), ip(bt.comp), dp(split), max(bt.n_comp, 1),
            float(mol_data.Gamma), float(mol_data.k), float(Gamma_eff_factor),
            float(HBAR_CODATA if hbar is None else hbar), int(kind))
    L = _lib.lib()
    n = int(L.mm_rate_spec_source_tables(*args, None, 0))
    if n == -2:
        return None
    if n <= 0:
        _lib.check(n, "mm_rate_spec_source_tables")
    buf = ctypes.create_string_buffer(n)
    L.mm_rate_spec_source_tables(*args, buf, n)
    return buf.value.decode()


def compile_source(src: str, arch: str = "gfx950") -> bytes:
    """hipRTC code object of a generated source (the library's flags)."""
    L = _lib.lib()
    n = int(L.mm_rtc_compile(src.encode(), arch.encode(), None, 0))
    if n <= 0:
        _lib.check(n, "mm_rtc_compile")
    buf = ctypes.create_string_buffer(n)
    L.mm_rtc_compile(src.encode(), arch.encode(), buf, n)
    return buf.raw[:n]


def _objdump():
    for cand in (shutil.which("llvm-objdump"), "/opt/rocm/llvm/bin/llvm-objdump"):
        if cand and os.path.exists(cand):
            return cand
    return None


def fp64_counts(code: bytes, kernel: str, arch: str = "gfx950"):
    """FP64 VALU instructions in `kernel` of a code object, by mnemonic class:
    {"fma", "mul", "add", "rcp", "flop"} with flop = 2 fma + mul + add.  The
    specialised step loop is straight-line code, so for the trajectory
    kernels these are the instructions of one step (plus a few outside the
    loop).  None when llvm-objdump is not available."""
    exe = _objdump()
    if exe is None:
        return None
    with tempfile.NamedTemporaryFile(suffix=".co", delete=False) as f:
        f.write(code)
        path = f.name
    try:
        out = subprocess.run([exe, "-d", f"--mcpu={arch}", path], capture_output=True, text=True,
                             check=True).stdout
    finally:
        os.unlink(path)
    body, inside = [], False
    for line in out.splitlines():
        m = re.match(r"^[0-9a-f]+ <(.+)>:", line)
        if m:
            inside = m.group(1) == kernel
            continue
        if inside and line.strip():
            body.append(line.split()[0])
    if not body:
        raise ValueError(f"kernel {kernel} not found in the code object")
    c = {"fma": 0, "mul": 0, "add": 0, "rcp": 0}
    for op in body:
        if op.startswith(("v_fma_f64", "v_fmac_f64")):
            c["fma"] += 1
        elif op.startswith("v_mul_f64"):
            c["mul"] += 1
        elif op.startswith("v_add_f64"):
            c["add"] += 1
        elif op.startswith("v_rcp_f64"):
            c["rcp"] += 1
    c["flop"] = 2 * c["fma"] + c["mul"] + c["add"]
    c["valu"] = sum(1 for op in body if op.startswith("v_"))
    return c
