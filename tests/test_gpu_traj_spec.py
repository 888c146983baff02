"""Plan-specialised 1D trajectory kernels (mm_rate_plan_specialize): the
plan's tables compiled into the code with hipRTC.  They must give the
generic kernels' trajectories bit for bit (the operations they leave out
are exact zeros), on both schedules, with and without recoil kicks, on
ragged ensembles with escapes, through the no-allocation entry point, for
every molecule whose plan admits them, and meet the oracle's trajectory bar
directly."""
import numpy as np
import pytest

import oracle
from molmot_amd import configs
from molmot_amd.constants import hbar as HBAR
from molmot_amd.obe.fields import beam_table

pytestmark = pytest.mark.gpu

DT = 1e-5
B = configs.B_GRAD_SI


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


@pytest.fixture(scope="module")
def plans(mol):
    """(generic, specialised) plans of the headline workload: SrOH in the
    four-frequency red MOT."""
    from molmot_amd.hip import HIPMolData
    beams = configs.red_mot_beams(mol)
    gen = HIPMolData(mol, beams, 1.0, hbar=HBAR, specialize=False)
    spec = HIPMolData(mol, beams, 1.0, hbar=HBAR, specialize=True)
    assert spec.specialized and not gen.specialized
    return gen, spec


def _ics(mol, n, seed, dev):
    import torch
    z0, v0 = configs.ensemble_initial_conditions(mol, n, seed=seed)
    return z0, v0, torch.from_numpy(z0).to(dev), torch.from_numpy(v0).to(dev)


def _same(a, b):
    import torch
    return all(torch.equal(x, y) for x, y in zip(a[:3], b[:3]))


@pytest.mark.parametrize("n,steps,kicks,schedule", [(64 * 50 + 23, 3001, True, "queue"),
                                                    (64 * 50 + 23, 3001, True, "grid"),
                                                    (2000, 1600, False, "queue"),
                                                    (2000, 1600, False, "grid"),
                                                    (100_000, 700, True, "auto")])
def test_specialised_bit_identical(mol, dev, plans, n, steps, kicks, schedule):
    """Ragged last group, escapes inside chunks (the [-15, 15] m/s sweep),
    odd and even step counts, and the auto schedule at 1e5 particles."""
    from molmot_amd.hip import recoil_kick_scale, traj_euler_device
    gen, spec = plans
    _, _, z0, v0 = _ics(mol, n, 17, dev)
    ks = recoil_kick_scale(mol) if kicks else 0.0
    print(f"specialised vs generic: n={n} steps={steps} kicks={kicks} schedule={schedule}", flush=True)
    g = traj_euler_device(gen, z0, v0, B, DT, steps, 0.015, kick_scale=ks, seed=21, schedule=schedule)
    s = traj_euler_device(spec, z0, v0, B, DT, steps, 0.015, kick_scale=ks, seed=21, schedule=schedule)
    na = g[2].cpu().numpy()
    assert (na < steps).any() and (na == steps).any()
    assert _same(g, s)


def test_specialised_vs_oracle(mol, dev, plans):
    """Direct oracle parity of the specialised kernels (no kicks): the
    oracle's escape steps, z to 1e-12 m and v to 1e-10 m/s, on 512 of the
    workload's particles over 3,000 steps."""
    from molmot_amd.hip import traj_euler_device
    _, spec = plans
    z0h, v0h, z0, v0 = _ics(mol, 512, 3, dev)
    print("specialised vs oracle: 512 x 3000", flush=True)
    zf, vf, na = (t.cpu().numpy() for t in traj_euler_device(spec, z0, v0, B, DT, 3000, 0.015)[:3])
    zo, vo, nao, _, _ = oracle.traj_euler(mol, beam_table(configs.red_mot_beams(mol)), z0h, v0h, DT, 3000, B,
                                          hbar=HBAR)
    assert np.array_equal(na, nao)
    assert np.max(np.abs(zf - zo)) < 1e-12
    assert np.max(np.abs(vf - vo)) < 1e-10


def test_specialised_noalloc(mol, dev, plans):
    """The no-allocation entry point runs the specialised kernels of a
    specialised plan: equal to the workspace entry point of the generic
    plan."""
    import torch
    from molmot_amd.hip import _lib, recoil_kick_scale, traj_euler_device
    gen, spec = plans
    _, _, z0, v0 = _ics(mol, 3000, 5, dev)
    ks = recoil_kick_scale(mol)
    g = traj_euler_device(gen, z0, v0, B, DT, 2000, 0.015, kick_scale=ks, seed=8, schedule="queue")
    zf = torch.empty_like(z0)
    vf = torch.empty_like(v0)
    na = torch.empty(3000, dtype=torch.int32, device=dev)
    rc = _lib.lib().mm_traj_euler_1d_noalloc(spec.handle, 3000, _lib.ptr(z0), _lib.ptr(v0), B, spec.mass, DT,
                                             2000, 0.015, _lib.ptr(zf), _lib.ptr(vf), _lib.ptr(na), ks, 8, 0, 0,
                                             2, _lib.stream_handle())
    _lib.check(rc, "mm_traj_euler_1d_noalloc")
    torch.cuda.synchronize()
    assert _same(g, (zf, vf, na))


def test_specialised_source_covers_only_coupled_pairs(mol, plans):
    """The generated evaluation holds one block per non-empty (ground,
    excited) pair (24 of SrOH's 48 in the red MOT) and one per ground
    state."""
    _, spec = plans
    src = spec.spec_source()
    assert src.count("const double om = ") == 24
    assert src.count("// ground state ") == mol.n_ground
    assert "mms_traj_queue_k1" in src and "mms_traj_grid_k0" in src


def test_built_molecules_specialise(dev):
    """The built CaOH and CaF tables (bundled structure) in the red MOT admit
    the specialised kernels too: bit-identical to their generic plans, kicks
    on, queue schedule."""
    import torch
    from molmot_amd.hip import HIPMolData, traj_euler_device
    from molmot_amd.molecules.builders import load_caf_from_sroh_structure, load_caoh_from_sroh_structure
    for load in (load_caoh_from_sroh_structure, load_caf_from_sroh_structure):
        m = load()
        beams = configs.red_mot_beams(m)
        gen = HIPMolData(m, beams, 1.0, hbar=HBAR, specialize=False)
        if not gen.specializable():
            pytest.fail(f"{load.__name__}: plan not specialisable")
        spec = HIPMolData(m, beams, 1.0, hbar=HBAR, specialize=True)
        rng = np.random.default_rng(2)
        z0 = torch.from_numpy(rng.normal(0, 1e-3, 1500)).to(dev)
        v0 = torch.from_numpy(rng.uniform(-15, 15, 1500)).to(dev)
        print(f"built molecule {load.__name__}", flush=True)
        g = traj_euler_device(gen, z0, v0, B, DT, 1800, 0.015, kick_scale=1e-3, seed=4, schedule="queue")
        s = traj_euler_device(spec, z0, v0, B, DT, 1800, 0.015, kick_scale=1e-3, seed=4, schedule="queue")
        assert _same(g, s), load.__name__


def test_unsupported_plans_and_release(mol, dev):
    """A plan outside the packed form (one sigma+ beam: pairs of one term)
    is refused with MM_EUNSUPPORTED and keeps the generic kernels;
    releasing a specialisation returns a plan to the generic kernels with
    the same results; the "auto" rule specialises a plan on its first
    launch above the work threshold."""
    import torch
    from molmot_amd.hip import HIPMolData, MolmotHIPError, traj_euler_device
    from molmot_amd.hip import plan as planmod
    from molmot_amd.obe.fields import LaserBeam
    p = HIPMolData(mol, beam_table([LaserBeam(1.0, mol.omega_J32 - 3e6, 2, 3.0)]), 1.0, hbar=HBAR)
    assert not p.specializable()
    with pytest.raises(MolmotHIPError, match="not packed"):
        p.specialize()
    assert not p.specialized
    q = HIPMolData(mol, configs.red_mot_beams(mol), 1.0, hbar=HBAR, specialize=True)
    _, _, z0, v0 = _ics(mol, 700, 9, dev)
    a = traj_euler_device(q, z0, v0, B, DT, 900, 0.015, kick_scale=1e-3, seed=2)
    q.specialize(False)
    assert not q.specialized
    b = traj_euler_device(q, z0, v0, B, DT, 900, 0.015, kick_scale=1e-3, seed=2)
    assert _same(a, b)
    r = HIPMolData(mol, configs.red_mot_beams(mol), 1.0, hbar=HBAR)
    old = planmod.SPECIALIZE_MIN_WORK
    try:
        planmod.SPECIALIZE_MIN_WORK = 700 * 900
        traj_euler_device(r, z0, v0, B, DT, 899, 0.015)
        assert not r.specialized
        c = traj_euler_device(r, z0, v0, B, DT, 900, 0.015, kick_scale=1e-3, seed=2)
        assert r.specialized
    finally:
        planmod.SPECIALIZE_MIN_WORK = old
    assert _same(a, c)
    torch.cuda.synchronize()
