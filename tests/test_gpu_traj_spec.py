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
    state, in each of its two force evaluators (without and with the
    populations)."""
    _, spec = plans
    src = spec.spec_source()
    assert src.count("const double om = ") == 2 * 24
    assert src.count("// ground state ") == 2 * mol.n_ground
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
    old = dict(planmod.SPECIALIZE_MIN_WORK)
    try:
        planmod.SPECIALIZE_MIN_WORK[planmod.FORWARD] = 700 * 900
        traj_euler_device(r, z0, v0, B, DT, 899, 0.015)
        assert not r.specialized
        c = traj_euler_device(r, z0, v0, B, DT, 900, 0.015, kick_scale=1e-3, seed=2)
        assert r.specialized == planmod.FORWARD
    finally:
        planmod.SPECIALIZE_MIN_WORK.update(old)
    assert _same(a, c)
    torch.cuda.synchronize()


# ---------------------------------------------------------------------------
# forward sensitivities (GRADIENT kind)
# ---------------------------------------------------------------------------

def _force_gradient_specialisation():
    """Context: every trajectory_tangent call specialises its plan."""
    import contextlib

    from molmot_amd.hip import plan as planmod

    @contextlib.contextmanager
    def ctx():
        old = dict(planmod.SPECIALIZE_MIN_WORK)
        planmod.SPECIALIZE_MIN_WORK[planmod.GRADIENT] = 0
        try:
            yield
        finally:
            planmod.SPECIALIZE_MIN_WORK.update(old)
    return ctx()


def test_specialised_tangent_vs_generic(mol, dev):
    """The specialised forward sensitivities against the generic kernels on
    the red-MOT workload (2,000 particles x 1,500 updates, escapes inside
    chunks), both schedules: the same force per step, so the same
    trajectories and update counts bit for bit; the sensitivities, built
    from a rearranged reverse pass, agree to 1e-9 of each column's scale."""
    import torch
    from molmot_amd.adjoint import ForceProblem, trajectory_tangent
    from molmot_amd.hip.plan import GRADIENT
    prob = ForceProblem(mol, configs.red_mot_beams(mol))
    theta = prob.theta0(B).numpy()
    assert prob.plan(theta).specializable(GRADIENT)
    _, _, z0, v0 = _ics(mol, 2000, 23, dev)
    for schedule in ("queue", "grid"):
        print(f"specialised tangent vs generic: 2000 x 1500, {schedule}", flush=True)
        g = trajectory_tangent(prob, theta, z0, v0, mol.mass, DT, 1500, 0.015, None, schedule=schedule)
        with _force_gradient_specialisation():
            s = trajectory_tangent(prob, theta, z0, v0, mol.mass, DT, 1500, 0.015, None, schedule=schedule)
        for k in ("z_final", "v_final", "n_done"):
            assert torch.equal(g[k], s[k]), k
        nd = g["n_done"].cpu().numpy()
        assert (nd < 1500).any() and (nd == 1500).any()
        a, b = g["sens"].cpu().numpy(), s["sens"].cpu().numpy()
        scale = np.max(np.abs(a), axis=0)
        assert np.all(np.abs(a - b) <= 1e-9 * scale + 1e-300), np.max(np.abs(a - b) / (scale + 1e-300))


def test_specialised_tangent_vs_torch_oracle(mol, dev):
    """Directly against the torch-autograd oracle (the CPU restatement of the
    force, differentiated through the Euler loop) on the red MOT: four
    particles, 300 updates, one escaping; gradient of a weighted sum of the
    final states w.r.t. the 7 parameters and the initial state to 1e-8."""
    import torch
    from oracle import torch_rate_eq
    from molmot_amd.adjoint import ForceProblem, trajectory_tangent, vjp_from_sens
    prob = ForceProblem(mol, configs.red_mot_beams(mol))
    theta = prob.theta0(B).numpy()
    Z0 = np.array([3e-3, -1e-3, 0.5e-3, 5e-3])
    V0 = np.array([-4.0, 2.0, 0.3, 14.0])
    CZ = np.array([1.0, -0.5, 2.0, 0.7])
    CV = np.array([1e-4, 3e-4, -2e-4, 1e-4])
    NU, ZESC = 300, 8e-3
    with _force_gradient_specialisation():
        o = trajectory_tangent(prob, theta, torch.tensor(Z0, device=dev), torch.tensor(V0, device=dev), mol.mass,
                               DT, NU, ZESC, None, schedule="grid")
    th = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    z0 = torch.tensor(Z0, requires_grad=True)
    v0 = torch.tensor(V0, requires_grad=True)
    bt = beam_table(prob.beams)
    z, v = z0, v0
    alive = torch.ones(4, dtype=torch.bool)
    done = np.zeros(4, dtype=int)
    for _ in range(NU):
        F = torch_rate_eq.force(mol, bt, th, v, z, hbar=HBAR)
        vn = v + F / mol.mass * DT
        zn = z + vn * DT
        v = torch.where(alive, vn, v)
        z = torch.where(alive, zn, z)
        done += alive.numpy()
        alive = alive & (z.abs() <= ZESC)
    (torch.tensor(CZ) * z + torch.tensor(CV) * v).sum().backward()
    nd = o["n_done"].cpu().numpy()
    assert np.array_equal(nd, done) and (nd < NU).any() and (nd == NU).any()
    rel = lambda a, b: np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)  # noqa: E731
    assert rel(o["z_final"].cpu().numpy(), z.detach().numpy()) < 1e-12
    g, g_z0, g_v0 = vjp_from_sens(o["sens"], torch.tensor(CZ, device=dev), torch.tensor(CV, device=dev))
    g = g.sum(0).cpu().numpy()
    gth = th.grad.numpy()
    for j in range(7):
        assert abs(g[j] - gth[j]) <= 1e-8 * max(abs(gth[j]), 1e-300), (j, g[j], gth[j])
    assert rel(g_z0.cpu().numpy(), z0.grad.numpy()) < 1e-8 and rel(g_v0.cpu().numpy(), v0.grad.numpy()) < 1e-8


def test_specialised_checkpointed_adjoint(mol, dev):
    """The checkpointed adjoint on plan-specialised kernels (forward pass with
    checkpoints, and the segments' force Jacobians): the same final states
    and update counts as the generic kernels bit for bit, the parameter and
    initial-state gradients to 1e-10 of their scale, and the forward
    sensitivities' gradient to 1e-9 (red MOT, 1,000 particles, 1,200 updates,
    checkpoints every 100)."""
    import torch
    from molmot_amd.adjoint import ForceProblem, trajectory_tangent, trajectory_vjp, vjp_from_sens
    from molmot_amd.hip import plan as planmod
    prob = ForceProblem(mol, configs.red_mot_beams(mol))
    theta = prob.theta0(B).numpy()
    _, _, z0, v0 = _ics(mol, 1000, 31, dev)
    cz = torch.ones(1000, dtype=torch.float64, device=dev)
    cv = torch.full((1000,), 1e-4, dtype=torch.float64, device=dev)
    print("specialised checkpointed adjoint: 1000 x 1200", flush=True)
    g = trajectory_vjp(prob, theta, z0, v0, mol.mass, DT, 1200, cz, cv, 0.015, None, 100)
    old = dict(planmod.SPECIALIZE_MIN_WORK)
    try:
        planmod.SPECIALIZE_MIN_WORK.update({planmod.FORWARD: 0, planmod.GRADIENT: 0})
        s = trajectory_vjp(prob, theta, z0, v0, mol.mass, DT, 1200, cz, cv, 0.015, None, 100)
        t = trajectory_tangent(prob, theta, z0, v0, mol.mass, DT, 1200, 0.015, None)
    finally:
        planmod.SPECIALIZE_MIN_WORK.update(old)
    for k in ("z_final", "v_final", "n_done"):
        assert torch.equal(g[k], s[k]), k
    nd = g["n_done"].cpu().numpy()
    assert (nd < 1200).any() and (nd == 1200).any()
    for k in ("g_theta", "g_z0", "g_v0"):
        a, b = g[k].cpu().numpy(), s[k].cpu().numpy()
        assert np.max(np.abs(a - b)) <= 1e-10 * np.max(np.abs(a)), k
    gt, gz, gv = vjp_from_sens(t["sens"], cz, cv)
    a, b = s["g_theta"].sum(0).cpu().numpy(), gt.sum(0).cpu().numpy()
    assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(a))
    assert torch.allclose(gz, s["g_z0"], rtol=1e-9, atol=0) and torch.allclose(gv, s["g_v0"], rtol=1e-9, atol=0)


# ---------------------------------------------------------------------------
# force map, scattered points and the force gradient at points
# ---------------------------------------------------------------------------

def test_specialised_map_and_points(mol, dev, plans):
    """The force map and scattered points of a specialised plan: F, R and the
    12 populations bit-identical to the generic kernels' (256 x 256 map over
    the config-2 ranges, with and without populations; 3,001 scattered
    points)."""
    import torch
    from molmot_amd.hip import rate_eq_map_device, rate_eq_points_device
    gen, spec = plans
    v = torch.linspace(-30, 30, 256, dtype=torch.float64, device=dev)
    z = torch.linspace(-10e-3, 10e-3, 256, dtype=torch.float64, device=dev)
    g = rate_eq_map_device(gen, v, z, B)
    s = rate_eq_map_device(spec, v, z, B)
    assert all(torch.equal(a, b) for a, b in zip(g, s))
    rng = np.random.default_rng(4)
    vp = torch.from_numpy(rng.uniform(-30, 30, 3001)).to(dev)
    zp = torch.from_numpy(rng.uniform(-10e-3, 10e-3, 3001)).to(dev)
    gp = rate_eq_points_device(gen, vp, zp, B)
    sp = rate_eq_points_device(spec, vp, zp, B)
    assert all(torch.equal(a, b) for a, b in zip(gp, sp))


def test_specialised_force_jacobian(mol, dev):
    """F and dF/d[4 detunings, split, s0, B', z, v] from the specialised
    gradient kernel: against the torch-autograd oracle (F to 1e-12, every
    column to 1e-8) and against the generic kernel (F bit for bit, columns to
    1e-12), red MOT, 400 points over the config-2 ranges."""
    import torch
    from oracle import torch_rate_eq
    from molmot_amd.adjoint import ForceProblem, force_and_jacobian
    prob = ForceProblem(mol, configs.red_mot_beams(mol))
    theta = prob.theta0(B).numpy()
    rng = np.random.default_rng(11)
    v = rng.uniform(-30, 30, 400)
    z = rng.uniform(-10e-3, 10e-3, 400)
    vt, zt = torch.from_numpy(v).to(dev), torch.from_numpy(z).to(dev)
    Fg, Jg = force_and_jacobian(prob, theta, vt, zt)
    Fs, Js = force_and_jacobian(prob, theta, vt, zt, specialize=True)
    assert torch.equal(Fg, Fs)
    colrel = lambda a, b: np.max(np.abs(a - b), axis=0) / np.maximum(np.max(np.abs(b), axis=0), 1e-300)  # noqa
    assert np.all(colrel(Js.cpu().numpy(), Jg.cpu().numpy()) < 1e-12)
    Jt, Ft = torch_rate_eq.jacobian(mol, beam_table(prob.beams), theta, v, z, hbar=HBAR)
    assert np.max(np.abs(Fs.cpu().numpy() - Ft)) / np.max(np.abs(Ft)) < 1e-12
    assert np.all(colrel(Js.cpu().numpy(), Jt) < 1e-8)


@pytest.mark.parametrize("n,steps,schedule", [(64 * 40 + 17, 2001, 1), (64 * 40 + 17, 2001, 2), (20_000, 1200, 0)])
def test_specialised_mc_temperature(mol, dev, plans, n, steps, schedule):
    """Monte-Carlo temperature sums (sum of v^2 and sample counts after i_eq,
    mm_traj_vsq_1d_ws) from the specialised kernels equal the generic
    kernels' bit for bit: one thread per particle, work queue, auto."""
    import torch
    from molmot_amd.hip import _lib
    gen, spec = plans
    _, _, z0, v0 = _ics(mol, n, 11, dev)
    L = _lib.lib()
    ws_bytes = int(L.mm_traj_vsq_workspace_bytes(n, steps))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    out = []
    for p in (gen, spec):
        vsq = torch.empty(n, dtype=torch.float64, device=dev)
        cnt = torch.empty(n, dtype=torch.int64, device=dev)
        _lib.check(L.mm_traj_vsq_1d_ws(p.handle, n, _lib.ptr(z0), _lib.ptr(v0), B, mol.mass, DT, steps, steps // 2,
                                       0.015, _lib.ptr(vsq), _lib.ptr(cnt), schedule, _lib.ptr(ws), ws_bytes,
                                       _lib.stream_handle()))
        torch.cuda.synchronize()
        out.append((vsq, cnt))
    assert int(out[1][1].min()) >= 0
    assert 0 < int(out[1][1].sum()) < n * (steps - steps // 2)   # some escape, some stay
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])


@pytest.mark.parametrize("n,steps,schedule", [(1000, 2000, 1), (1000, 2000, 2), (100_000, 600, 0)])
def test_specialised_axis3d(mol, dev, plans, n, steps, schedule):
    """Per-axis 3D model (mm_traj_axis3d_ws, lane quads) from the specialised
    kernels equals the generic kernels' bit for bit: one quad per particle,
    work queue, auto at 1e5 particles; a launch that saves the trajectory
    (generic kernel on the specialised plan) ends at the same state."""
    import torch
    from molmot_amd.hip import _lib
    gen, spec = plans
    L = _lib.lib()
    r0h, v0h = configs.ensemble_initial_conditions_3d(mol, n, seed=61)
    r0, v0 = torch.from_numpy(r0h).to(dev), torch.from_numpy(v0h).to(dev)

    def run(p, save=False):
        ws_bytes = int(L.mm_traj_axis3d_workspace_bytes(n, steps))
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        rf, vf = torch.empty_like(r0), torch.empty_like(v0)
        na = torch.empty(n, dtype=torch.int32, device=dev)
        rt = torch.empty((n, steps, 3), dtype=torch.float64, device=dev) if save else None
        vt = torch.empty_like(rt) if save else None
        _lib.check(L.mm_traj_axis3d_ws(p.handle, n, _lib.ptr(r0), _lib.ptr(v0), B, mol.mass, DT, steps, 0.015,
                                       _lib.ptr(rt), _lib.ptr(vt), _lib.ptr(rf), _lib.ptr(vf), _lib.ptr(na),
                                       1 if save else schedule, _lib.ptr(ws), ws_bytes, _lib.stream_handle()))
        torch.cuda.synchronize()
        return rf, vf, na

    g, s = run(gen), run(spec)
    na = g[2].cpu().numpy()
    assert (na < steps).any() and (na == steps).any()
    assert all(torch.equal(x, y) for x, y in zip(g, s))
    if n <= 1000:
        assert all(torch.equal(x, y) for x, y in zip(g, run(spec, save=True)))


@pytest.mark.parametrize("kicks", [False, True])
def test_specialised_ensemble_statistics(mol, dev, plans, kicks):
    """Ensemble moments at sample steps (mm_traj_stats_1d: count, mean and
    M2 of z and v of the molecules inside r_alive) from the specialised
    kernel equal the generic kernel's bit for bit, recoil kicks on and off,
    on a ragged ensemble with escapes."""
    import torch
    from molmot_amd.hip import _lib, recoil_kick_scale
    gen, spec = plans
    n, steps, bins = 64 * 70 + 9, 3001, 25
    _, _, z0, v0 = _ics(mol, n, 5, dev)
    samp = torch.from_numpy(np.linspace(0, steps - 1, bins).astype(np.int32)).to(dev)
    ks = recoil_kick_scale(mol) if kicks else 0.0
    L = _lib.lib()
    wb = int(L.mm_traj_stats_workspace_bytes(n, bins))
    ws = torch.empty(max(wb, 8), dtype=torch.uint8, device=dev)
    out = []
    for p in (gen, spec):
        o = torch.empty((bins, 5), dtype=torch.float64, device=dev)
        _lib.check(L.mm_traj_stats_1d(p.handle, n, _lib.ptr(z0), _lib.ptr(v0), B, mol.mass, DT, steps, 0.015, 0.015,
                                      _lib.ptr(samp), bins, _lib.ptr(o), ks, 7, 0, 0, _lib.ptr(ws), wb,
                                      _lib.stream_handle()))
        torch.cuda.synchronize()
        out.append(o)
    alive = out[1][:, 0].cpu().numpy()
    assert alive[0] == n and 0 < alive[-1] < n
    assert torch.equal(out[0], out[1])


@pytest.mark.parametrize("kicks", [False, True])
def test_specialised_saved_trajectories(mol, dev, plans, kicks):
    """Launches that save every step (simulate_trajectory's path) run the
    specialised saving kernel: z(t), v(t) and the finals equal the generic
    kernel's bit for bit, escapes included (states after an escape repeat
    the escape state)."""
    import torch
    from molmot_amd.hip import recoil_kick_scale, traj_euler_device
    gen, spec = plans
    _, _, z0, v0 = _ics(mol, 300, 41, dev)
    ks = recoil_kick_scale(mol) if kicks else 0.0
    g = traj_euler_device(gen, z0, v0, B, DT, 1200, 0.015, save_full=True, kick_scale=ks, seed=9)
    s = traj_euler_device(spec, z0, v0, B, DT, 1200, 0.015, save_full=True, kick_scale=ks, seed=9)
    torch.cuda.synchronize()
    na = g[2].cpu().numpy()
    assert (na < 1200).any() and (na == 1200).any()
    assert all(torch.equal(x, y) for x, y in zip(g, s))
