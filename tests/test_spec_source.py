"""The generator of the plan-specialised trajectory kernels, on the host (no
device): the generated source for SrOH in the four-frequency red MOT holds
one block per coupled (ground, excited) pair with the plan's values as exact
literals, it compiles with hipRTC for gfx950, and its step executes far fewer
FP64 operations than the reference algorithm's dense count."""
import re
import struct

import numpy as np
import pytest

from molmot_amd import configs
from molmot_amd.constants import hbar as HBAR
from molmot_amd.obe.fields import LaserBeam, beam_table


def _lib_or_skip():
    from molmot_amd.hip import _lib
    if not _lib.available():
        pytest.skip("libmolmot_hip.so not built")


@pytest.fixture(scope="module")
def sroh_source(mol):
    _lib_or_skip()
    from molmot_amd.hip import spec
    src = spec.spec_source(mol, configs.red_mot_beams(mol), hbar=HBAR)
    assert src is not None
    return src


def test_source_structure(mol, sroh_source):
    """24 of SrOH's 48 pairs are driven by sigma beams in the red MOT (each by
    two +z and two -z beams): one block each in each of the two force
    evaluators (force only, and force with populations for the map); every
    ground state has one."""
    d2 = np.asarray(mol.d_squared)
    coupled = int(np.sum((d2[:, :, 0] >= 1e-15) | (d2[:, :, 2] >= 1e-15)))
    assert coupled == 24
    fast, fast_pop = sroh_source.split("bool mms_fast_pop(")
    fast = fast.split("bool mms_fast(")[1]
    for body in (fast, fast_pop):
        body = body.split("__global__")[0]
        assert body.count("const double om = ") == coupled
        assert body.count("// ground state ") == mol.n_ground
    for k in ("mms_traj_queue_k0", "mms_traj_queue_k1", "mms_traj_grid_k0", "mms_traj_grid_k1",
              "mms_rate_map", "mms_rate_points", "mms_vsq_grid", "mms_vsq_queue", "mms_axis3d_grid",
              "mms_axis3d_queue", "mms_traj_stats_k0", "mms_traj_stats_k1", "mms_traj_save_k0", "mms_traj_save_k1"):
        assert k in sroh_source


def test_literals_are_exact(mol, sroh_source):
    """Every beam frequency and ground energy appears as an exact hexadecimal
    literal (no decimal rounding of the tables)."""
    bt = beam_table(configs.red_mot_beams(mol))
    lits = {float.fromhex(m) for m in re.findall(r"\(([-+]?0x[0-9a-fA-Fp.+-]+)\)", sroh_source)}
    for f in bt.freqs:
        assert float(f) in lits
    for e in np.asarray(mol.energies)[:mol.n_ground]:
        assert float(e) in lits
    # bitwise, not just equal: the literal reproduces the double
    f0 = float(bt.freqs[0])
    assert struct.pack("<d", float.fromhex(f0.hex())) == struct.pack("<d", f0)


def test_compiles_for_gfx950(sroh_source):
    """hipRTC builds the four kernels for gfx950 without a device; the step
    executes ~1,300 FP64 flop (2 x FMA + MUL + ADD), against the reference
    algorithm's dense 4,272."""
    from molmot_amd.hip import spec
    code = spec.compile_source(sroh_source)
    assert code[:4] == b"\x7fELF"
    for k in (b"mms_traj_queue_k1", b"mms_traj_grid_k0"):
        assert k in code
    c1 = spec.fp64_counts(code, "mms_traj_queue_k1")
    if c1 is None:
        pytest.skip("llvm-objdump not available")
    c0 = spec.fp64_counts(code, "mms_traj_queue_k0")
    assert 800 < c1["flop"] < 2000
    assert c0["flop"] <= c1["flop"]
    # at most one reciprocal per coupled pair and per ground state, three in
    # the excited chain and one for the normalisation (identical pairs of
    # degenerate ground states share theirs)
    assert 12 <= c1["rcp"] <= 24 + 12 + 3 + 1


def test_unsupported_tables(mol):
    """A plan outside the packed form (one sigma+ beam) has no specialised
    source; the built CaOH table in the red MOT has one."""
    _lib_or_skip()
    from molmot_amd.hip import spec
    from molmot_amd.molecules.builders import load_caoh_from_sroh_structure
    assert spec.spec_source(mol, beam_table([LaserBeam(1.0, mol.omega_J32 - 3e6, 2, 3.0)])) is None
    m = load_caoh_from_sroh_structure()
    assert spec.spec_source(m, configs.red_mot_beams(m), hbar=HBAR) is not None


def test_compile_errors_are_reported():
    """A source that does not compile returns MM_EHIP with the compiler log."""
    _lib_or_skip()
    from molmot_amd.hip import MolmotHIPError, spec
    with pytest.raises(MolmotHIPError, match="hipRTC compile failed"):
        spec.compile_source("this is not C++")


def test_gradient_source_compiles(mol):
    """The gradient kind (forward sensitivities, checkpointed-adjoint
    segments, force-gradient points) for SrOH in the red MOT: one reverse
    block per coupled pair as well, compiled for gfx950 with its four
    kernels; it executes ~2.6x the forward step's FP64 work."""
    _lib_or_skip()
    from molmot_amd.hip import spec
    src = spec.spec_source(mol, configs.red_mot_beams(mol), hbar=HBAR, kind=2)
    assert src is not None
    code = spec.compile_source(src)
    for k in (b"mms_tangent_queue", b"mms_tangent_grid", b"mms_vjp_segment", b"mms_rate_grad_points"):
        assert k in code
    c = spec.fp64_counts(code, "mms_tangent_queue")
    if c is None:
        pytest.skip("llvm-objdump not available")
    f = spec.fp64_counts(spec.compile_source(spec.spec_source(mol, configs.red_mot_beams(mol), hbar=HBAR)),
                         "mms_traj_queue_k1")
    assert 2.0 < c["flop"] / f["flop"] < 3.5
