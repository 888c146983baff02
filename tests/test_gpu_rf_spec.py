"""Plan-specialised RF MOT kernels (mm_rate_plan_specialize_rf).  The two
switching-phase plans of SrOH's RF MOT (molmot/mot/simulator.py:49-90) are
outside the packed form: a pair is driven by one +z and one -z beam per
frequency.  The generated kernels evaluate both phases with the tables as
constants, in the general pair path's order, and give bit-identical
trajectories, escape steps and temperature sums to the generic RF kernels;
a second plan other than the one they were built with takes the generic
kernels."""
import numpy as np
import pytest

from molmot_amd.constants import hbar as HBAR
from molmot_amd.obe.fields import make_rf_mot_beams

pytestmark = pytest.mark.gpu

B_GRAD, DT, N_STEPS, Z_ESC = 0.16, 1e-5, 400, 0.015


@pytest.fixture(scope="module")
def plans(mol):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from molmot_amd.hip.plan import HIPMolData
    return tuple(HIPMolData(mol, make_rf_mot_beams(mol, delta_Gamma=-1.0, s0=1.0, phase=ph), hbar=HBAR,
                            specialize=False) for ph in (0, 1))


def _ensemble(n, seed):
    import torch
    rng = np.random.default_rng(seed)
    z0 = torch.tensor(rng.normal(0, 2e-3, n), dtype=torch.float64, device="cuda")
    v0 = torch.tensor(rng.normal(0, 3.0, n), dtype=torch.float64, device="cuda")
    return z0, v0


def test_rf_source(mol, plans):
    """The phase plans are not packed (no single-plan specialisation), but
    the pair admits the RF kernels: two evaluations, five kernels."""
    p0, p1 = plans
    assert not p0.specializable() and not p1.specializable()
    assert p0.rf_specializable(p1)
    src = p0.spec_source_rf(p1)
    for k in ("mms_fast(", "mms_fast_b(", "mms_rf_traj_k0", "mms_rf_traj_k1", "mms_rf_save_k0", "mms_rf_save_k1",
              "mms_rf_vsq"):
        assert k in src
    assert src.count("// ground state ") == 2 * mol.n_ground


@pytest.mark.parametrize("kick_scale,save", [(0.0, False), (0.02, False), (0.0, True), (0.02, True)])
def test_rf_trajectories_bit_identical(plans, kick_scale, save):
    import torch
    from molmot_amd.hip.trajectories import traj_euler_rf_device
    p0, p1 = plans
    z0, v0 = _ensemble(1003, 11)
    kw = dict(save_full=save, kick_scale=kick_scale, seed=5, stream_id=2)
    p0.specialize_rf(None)
    assert not p0.specialized_rf(p1)
    ref = traj_euler_rf_device(p0, p1, z0, v0, B_GRAD, DT, N_STEPS, Z_ESC, **kw)
    p0.specialize_rf(p1)
    assert p0.specialized_rf(p1) and not p1.specialized_rf(p0)
    out = traj_euler_rf_device(p0, p1, z0, v0, B_GRAD, DT, N_STEPS, Z_ESC, **kw)
    torch.cuda.synchronize()
    na = ref[2].cpu().numpy()
    assert (na < N_STEPS).any() and (na == N_STEPS).any()
    for a, b in zip(ref, out):
        if a is None:
            assert b is None
        else:
            assert torch.equal(a, b)


def test_rf_temperature_sums_bit_identical(mol, plans):
    import torch
    from molmot_amd.hip import _lib
    p0, p1 = plans
    n, n_steps, i_eq = 2051, 600, 200
    z0, v0 = _ensemble(n, 3)

    def run():
        vsq = torch.empty(n, dtype=torch.float64, device="cuda")
        cnt = torch.empty(n, dtype=torch.int64, device="cuda")
        rc = _lib.lib().mm_traj_vsq_rf_1d(p0.handle, p1.handle, n, _lib.ptr(z0), _lib.ptr(v0), B_GRAD, p0.mass, DT,
                                          n_steps, i_eq, Z_ESC, _lib.ptr(vsq), _lib.ptr(cnt), _lib.stream_handle())
        _lib.check(rc, "mm_traj_vsq_rf_1d")
        torch.cuda.synchronize()
        return vsq, cnt

    p0.specialize_rf(None)
    ref = run()
    p0.specialize_rf(p1)
    out = run()
    assert torch.equal(ref[0], out[0]) and torch.equal(ref[1], out[1])
    c = ref[1].cpu().numpy()
    assert (c > 0).any() and (c < n_steps - i_eq).any()


def test_rf_kernels_belong_to_their_pair(mol, plans):
    """A plan built later (another serial) does not run p0's RF kernels."""
    from molmot_amd.hip.plan import HIPMolData
    p0, p1 = plans
    p0.specialize_rf(p1)
    other = HIPMolData(mol, make_rf_mot_beams(mol, delta_Gamma=-1.0, s0=1.0, phase=1), hbar=HBAR,
                       specialize=False)
    assert p0.specialized_rf(p1) and not p0.specialized_rf(other)
    other.close()
    p0.specialize_rf(None)
    assert not p0.specialized_rf(p1)
