"""The work-queue schedule of the 1D trajectory ensemble and of its forward
sensitivities (mm_traj_euler_1d_ws, mm_traj_tangent_1d_ws; schedule 2): a
persistent grid of wavefronts takes
(group of 64 particles, chunk of steps) units from a queue.  It must give
exactly what one thread per particle gives (same arithmetic per step, same
Philox counters), on ragged ensembles with escapes inside a chunk, with and
without recoil kicks, on packed and unpacked plans, under HIP-graph replay,
and meet the oracle's trajectory bar directly."""
import numpy as np
import pytest

import oracle
from molmot_amd import configs
from molmot_amd.constants import hbar as HBAR
from molmot_amd.obe.fields import beam_table, make_rf_mot_beams

pytestmark = pytest.mark.gpu

DT = 1e-5
B = 0.214


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


@pytest.fixture(scope="module")
def red_plan(mol):
    from molmot_amd.hip import HIPMolData
    return HIPMolData(mol, configs.red_mot_beams(mol), 1.0, hbar=HBAR)


def _ics(mol, n, seed, dev):
    import torch
    z0, v0 = configs.ensemble_initial_conditions(mol, n, seed=seed)
    return z0, v0, torch.from_numpy(z0).to(dev), torch.from_numpy(v0).to(dev)


def _same(a, b):
    import torch
    return all(torch.equal(x, y) for x, y in zip(a[:3], b[:3]))


@pytest.mark.parametrize("n,steps,kicks", [(1000, 1500, True), (1000, 1500, False), (64 * 40 + 17, 2600, True),
                                           (100_000, 600, True)])
def test_queue_bit_identical_to_grid(mol, dev, red_plan, n, steps, kicks):
    """Ragged last group (n not a multiple of 64), 2..20 chunks, more
    wavefronts than groups (waits on the queue) and, at 1e5 particles, the
    schedule that auto picks.  The [-15, 15] m/s sweep makes particles
    escape inside chunks."""
    import torch
    from molmot_amd.hip import _lib, recoil_kick_scale, traj_euler_device
    assert _lib.lib().mm_traj_euler_workspace_bytes(n, steps) > 0
    _, _, z0, v0 = _ics(mol, n, 11, dev)
    ks = recoil_kick_scale(mol) if kicks else 0.0
    print(f"queue vs grid: n={n} steps={steps} kicks={kicks}", flush=True)
    g = traj_euler_device(red_plan, z0, v0, B, DT, steps, 0.015, kick_scale=ks, seed=3, schedule="grid")
    q = traj_euler_device(red_plan, z0, v0, B, DT, steps, 0.015, kick_scale=ks, seed=3, schedule="queue")
    torch.cuda.synchronize()
    na = g[2].cpu().numpy()
    assert (na < steps).any() and (na == steps).any()   # escapes and survivors
    assert (q[2].cpu().numpy() >= 0).all()               # no timed-out wait
    assert _same(g, q)
    if n >= 100_000:
        a = traj_euler_device(red_plan, z0, v0, B, DT, steps, 0.015, kick_scale=ks, seed=3)
        assert _same(g, a)


def test_auto_schedule_choice(mol, dev, red_plan):
    """auto takes the queue where one thread per particle leaves a last
    round of wavefronts well short of full, the grid where it is balanced
    (on the 256-CU MI355X: 1,024 SIMDs, three 64-thread blocks of this
    kernel per SIMD)."""
    import torch
    from molmot_amd.hip import recoil_kick_scale, traj_euler_schedule
    ks = recoil_kick_scale(mol)
    if torch.cuda.get_device_properties(0).multi_processor_count != 256:
        pytest.skip("sizes below are for 256 CUs")
    assert traj_euler_schedule(red_plan, 100_000, 30_000, kick_scale=ks) == "queue"   # 1,563 groups
    assert traj_euler_schedule(red_plan, 400_000, 30_000, kick_scale=ks) == "queue"   # 6,250 groups
    assert traj_euler_schedule(red_plan, 131_072, 30_000, kick_scale=ks) == "grid"    # 2 per SIMD
    assert traj_euler_schedule(red_plan, 65_536, 30_000, kick_scale=ks) == "grid"     # 1 per SIMD
    assert traj_euler_schedule(red_plan, 100_000, 30_000, save_full=True) == "grid"
    assert traj_euler_schedule(red_plan, 100_000, 200) == "grid"                      # too short


def test_queue_vs_oracle_packed(mol, dev, red_plan):
    """The queue on the packed-plan kernel against the oracle: identical
    escape steps, end states to the ensemble bar (80% within 1e-12, all
    within 1e-10 m and 1e-8 m/s)."""
    from molmot_amd.hip import traj_euler_device
    z0h, v0h, z0, v0 = _ics(mol, 200, 23, dev)
    zf, vf, na, _, _ = traj_euler_device(red_plan, z0, v0, B, DT, 1000, 0.015, schedule="queue")
    zf, vf, na = zf.cpu().numpy(), vf.cpu().numpy(), na.cpu().numpy()
    zo, vo, nao, _, _ = oracle.traj_euler(mol, beam_table(configs.red_mot_beams(mol)), z0h, v0h, DT, 1000, B,
                                          hbar=HBAR)
    assert np.array_equal(na, nao)
    dz, dv = np.abs(zf - zo), np.abs(vf - vo)
    assert np.quantile(dz, 0.8) < 1e-12 and np.quantile(dv, 0.8) < 1e-12
    assert dz.max() < 1e-10 and dv.max() < 1e-8


def test_queue_unpacked_plan(mol, dev):
    """RF-MOT beams of one phase (a plan without packed records): queue =
    grid bit for bit, and the oracle's bar."""
    from molmot_amd.hip import HIPMolData, traj_euler_device
    bt = beam_table(make_rf_mot_beams(mol, -1.0, 2.0, phase=0))
    plan = HIPMolData(mol, bt, 1.0, hbar=HBAR)
    rng = np.random.default_rng(5)
    n = 300
    z0h, v0h = rng.normal(0, 1e-3, n), rng.uniform(-5, 5, n)
    import torch
    z0, v0 = torch.from_numpy(z0h).to(dev), torch.from_numpy(v0h).to(dev)
    g = traj_euler_device(plan, z0, v0, 0.16, DT, 1200, 0.015, schedule="grid")
    q = traj_euler_device(plan, z0, v0, 0.16, DT, 1200, 0.015, schedule="queue")
    assert _same(g, q)
    zo, vo, nao, _, _ = oracle.traj_euler(mol, bt, z0h, v0h, DT, 1200, 0.16, hbar=HBAR)
    zf, vf, na = (t.cpu().numpy() for t in q[:3])
    assert np.array_equal(na, nao)
    dz, dv = np.abs(zf - zo), np.abs(vf - vo)
    assert np.quantile(dz, 0.8) < 1e-12 and np.quantile(dv, 0.8) < 1e-12
    assert dz.max() < 1e-10 and dv.max() < 1e-8


def test_queue_graph_replay(mol, dev, red_plan):
    """The queue launch (clear kernel, persistent kernel, status kernel)
    captured into a HIP graph and replayed three times, with eager launches
    and synchronisations in between: every replay equals the eager result."""
    import torch
    from molmot_amd.hip import _lib, recoil_kick_scale, traj_euler_device
    from molmot_amd.hip.graphs import CapturedLaunches
    n, steps = 3000, 900
    L = _lib.lib()
    ws_bytes = int(L.mm_traj_euler_workspace_bytes(n, steps))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    _, _, z0, v0 = _ics(mol, n, 31, dev)
    zf = torch.empty(n, dtype=torch.float64, device=dev)
    vf = torch.empty_like(zf)
    na = torch.empty(n, dtype=torch.int32, device=dev)
    ks = recoil_kick_scale(mol)

    def launch():
        rc = L.mm_traj_euler_1d_ws(red_plan.handle, n, _lib.ptr(z0), _lib.ptr(v0), B, red_plan.mass, DT, steps,
                                   0.015, None, None, _lib.ptr(zf), _lib.ptr(vf), _lib.ptr(na), ks, 9, 0, 0, 2,
                                   _lib.ptr(ws), ws_bytes, _lib.stream_handle())
        _lib.check(rc, "mm_traj_euler_1d_ws")

    ref = traj_euler_device(red_plan, z0, v0, B, DT, steps, 0.015, kick_scale=ks, seed=9, schedule="grid")
    graph = CapturedLaunches(launch)
    for r in range(3):
        zf.fill_(0.0)
        na.fill_(0)
        other = traj_euler_device(red_plan, z0[:128], v0[:128], B, DT, 300, 0.015, schedule="queue")
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(zf, ref[0]) and torch.equal(vf, ref[1]) and torch.equal(na, ref[2]), r
        assert other[2].min() >= 0


def test_queue_rejects_what_it_cannot_run(mol, dev, red_plan):
    import torch
    from molmot_amd.hip import traj_euler_device
    from molmot_amd.hip._lib import MolmotHIPError
    _, _, z0, v0 = _ics(mol, 100, 1, dev)
    with pytest.raises(ValueError):
        traj_euler_device(red_plan, z0, v0, B, DT, 1000, schedule="fastest")
    with pytest.raises(MolmotHIPError, match="queue"):
        traj_euler_device(red_plan, z0, v0, B, DT, 1000, save_full=True, schedule="queue")
    with pytest.raises(MolmotHIPError, match="queue"):
        traj_euler_device(red_plan, z0, v0, B, DT, 200, schedule="queue")
    # auto falls back to the grid where the queue does not apply
    a = traj_euler_device(red_plan, z0, v0, B, DT, 200)
    g = traj_euler_device(red_plan, z0, v0, B, DT, 200, schedule="grid")
    assert _same(a, g)
    torch.cuda.synchronize()


def _tangent_pair(prob, theta, z0, v0, mass, n_updates, z_escape, n_up=None):
    import torch
    from molmot_amd.adjoint import trajectory_tangent
    g = trajectory_tangent(prob, theta, z0, v0, mass, DT, n_updates, z_escape, n_up, schedule="grid")
    q = trajectory_tangent(prob, theta, z0, v0, mass, DT, n_updates, z_escape, n_up, schedule="queue")
    torch.cuda.synchronize()
    return g, q


def _same_dict(g, q):
    import torch
    return all(torch.equal(g[k], q[k]) for k in ("z_final", "v_final", "n_done", "sens"))


def test_tangent_queue_bit_identical(mol, dev):
    """Forward sensitivities: queue = grid bit for bit (state, updates done
    and the 2 x 9 Jacobian) on the 4 particles of the adjoint tests (one
    escapes; 300 updates, 3 chunks) and on 2,577 red-MOT particles over 1,000
    updates with per-particle update limits (n_up, -1 = none) and escapes
    inside chunks."""
    import torch
    from molmot_amd.adjoint import ForceProblem
    from molmot_amd.obe.fields import make_dc_mot_beams
    prob = ForceProblem(mol, make_dc_mot_beams(mol, -0.20, 0.70, 1.0))
    theta = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.16])
    z0 = torch.tensor([3e-3, -1e-3, 0.5e-3, 5e-3], device=dev)
    v0 = torch.tensor([-4.0, 2.0, 0.3, 12.0], device=dev)
    g, q = _tangent_pair(prob, theta, z0, v0, mol.mass, 300, 8e-3)
    assert (g["n_done"].cpu().numpy() < 300).any()
    assert _same_dict(g, q)
    prob = ForceProblem(mol, configs.red_mot_beams(mol))
    theta = prob.theta0(B).numpy()
    n = 64 * 40 + 17
    _, _, z0, v0 = _ics(mol, n, 41, dev)
    rng = np.random.default_rng(3)
    n_up = torch.from_numpy(np.where(rng.random(n) < 0.3, rng.integers(0, 1000, n), -1).astype(np.int32)).to(dev)
    g, q = _tangent_pair(prob, theta, z0, v0, mol.mass, 1000, 0.015, n_up)
    nd = g["n_done"].cpu().numpy()
    assert (nd < 1000).any() and (nd == 1000).any() and (q["n_done"].cpu().numpy() >= 0).all()
    assert _same_dict(g, q)


def test_tangent_auto_at_1e5(mol, dev):
    """1e5 red-MOT particles over 600 updates: auto (the queue on a 256-CU
    device) = grid bit for bit."""
    import torch
    from molmot_amd.adjoint import ForceProblem, trajectory_tangent
    prob = ForceProblem(mol, configs.red_mot_beams(mol))
    theta = prob.theta0(B).numpy()
    _, _, z0, v0 = _ics(mol, 100_000, 43, dev)
    g = trajectory_tangent(prob, theta, z0, v0, mol.mass, DT, 600, 0.015, None, schedule="grid")
    a = trajectory_tangent(prob, theta, z0, v0, mol.mass, DT, 600, 0.015, None)
    torch.cuda.synchronize()
    assert (a["n_done"].cpu().numpy() >= 0).all()
    assert _same_dict(g, a)


def _vsq(plan, mass, z0, v0, n_steps, i_eq, schedule, dev):
    import torch
    from molmot_amd.hip import _lib
    L = _lib.lib()
    n = z0.numel()
    ws_bytes = int(L.mm_traj_vsq_workspace_bytes(n, n_steps))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    vsq = torch.empty(n, dtype=torch.float64, device=dev)
    cnt = torch.empty(n, dtype=torch.int64, device=dev)
    rc = L.mm_traj_vsq_1d_ws(plan.handle, n, _lib.ptr(z0), _lib.ptr(v0), B, mass, DT, n_steps, i_eq, 0.015,
                             _lib.ptr(vsq), _lib.ptr(cnt), schedule, _lib.ptr(ws), ws_bytes, _lib.stream_handle())
    _lib.check(rc, "mm_traj_vsq_1d_ws")
    torch.cuda.synchronize()
    return vsq, cnt


def test_vsq_queue_bit_identical_and_oracle(mol, dev, red_plan):
    """Monte-Carlo temperature statistics (mm_traj_vsq_1d_ws): queue = one
    thread per particle bit for bit on 2,577 particles x 2,600 steps (the
    sweep makes particles escape inside chunks) and at 1e5 particles under
    auto; and the queue against the oracle's traj_vsq on 200 particles x
    1,000 steps: identical sample counts, per-particle sums within 1e-8 of
    the largest and the total within 1e-10 relative (the trajectory bar:
    end states agree to ~1e-12, a few stalled particles to 1e-10)."""
    import torch
    n = 64 * 40 + 17
    _, _, z0, v0 = _ics(mol, n, 51, dev)
    g = _vsq(red_plan, mol.mass, z0, v0, 2600, 1300, 1, dev)
    q = _vsq(red_plan, mol.mass, z0, v0, 2600, 1300, 2, dev)
    c = g[1].cpu().numpy()
    assert (c < 1299).any() and (c == 1300).any() and (q[1].cpu().numpy() >= 0).all()
    assert torch.equal(g[0], q[0]) and torch.equal(g[1], q[1])
    _, _, z0, v0 = _ics(mol, 100_000, 52, dev)
    g = _vsq(red_plan, mol.mass, z0, v0, 600, 300, 1, dev)
    a = _vsq(red_plan, mol.mass, z0, v0, 600, 300, 0, dev)
    assert torch.equal(g[0], a[0]) and torch.equal(g[1], a[1])
    z0h, v0h, z0, v0 = _ics(mol, 200, 53, dev)
    vq, cq = _vsq(red_plan, mol.mass, z0, v0, 1000, 400, 2, dev)
    vo, co = oracle.traj_vsq(mol, beam_table(configs.red_mot_beams(mol)), z0h, v0h, DT, 1000, 400, B, hbar=HBAR)
    assert np.array_equal(cq.cpu().numpy(), co)
    vq = vq.cpu().numpy()
    assert np.max(np.abs(vq - vo)) <= 1e-8 * np.max(np.abs(vo))
    assert abs(vq.sum() - vo.sum()) <= 1e-10 * abs(vo.sum())


def test_axis3d_queue_bit_identical(mol, dev, red_plan):
    """Per-axis 3D model (mm_traj_axis3d_ws): queue (groups of 16 particles,
    lane quads) = one quad per particle bit for bit on 1,000 particles x
    2,000 steps (a ragged last group, particles leaving the 15 mm sphere
    inside chunks) and at 1e5 particles x 600 steps under auto."""
    import torch
    from molmot_amd.hip import _lib
    L = _lib.lib()

    def run(r0, v0, ns, schedule):
        n = r0.shape[0]
        ws_bytes = int(L.mm_traj_axis3d_workspace_bytes(n, ns))
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
        rf, vf = torch.empty_like(r0), torch.empty_like(v0)
        na = torch.empty(n, dtype=torch.int32, device=dev)
        rc = L.mm_traj_axis3d_ws(red_plan.handle, n, _lib.ptr(r0), _lib.ptr(v0), B, mol.mass, DT, ns, 0.015, None,
                                 None, _lib.ptr(rf), _lib.ptr(vf), _lib.ptr(na), schedule, _lib.ptr(ws), ws_bytes,
                                 _lib.stream_handle())
        _lib.check(rc, "mm_traj_axis3d_ws")
        torch.cuda.synchronize()
        return rf, vf, na

    for n, ns, sch in ((1000, 2000, 2), (100_000, 600, 0)):
        r0h, v0h = configs.ensemble_initial_conditions_3d(mol, n, seed=61)
        r0, v0 = torch.from_numpy(r0h).to(dev), torch.from_numpy(v0h).to(dev)
        g = run(r0, v0, ns, 1)
        q = run(r0, v0, ns, sch)
        na = g[2].cpu().numpy()
        assert (na < ns).any() and (na == ns).any() and (q[2].cpu().numpy() >= 0).all()
        assert all(torch.equal(x, y) for x, y in zip(g, q)), n


@pytest.mark.parametrize("ng,ne,seed", [(36, 4, 3), (4, 2, 5), (3, 1, 2), (2, 1, 7)])
def test_queue_other_plan_sizes(dev, ng, ne, seed):
    """The queue kernels of the other compiled sizes (synthetic ng + ne
    level structures, DC-MOT beams): queue = one thread per particle bit for
    bit on 700 particles (a ragged last group) x 1,200 steps with kicks."""
    import torch
    from molmot_amd.hip import HIPMolData, traj_euler_device
    from molmot_amd.molecules import synthetic_molecule
    from molmot_amd.obe.fields import make_dc_mot_beams
    m = synthetic_molecule(ng, ne, seed=seed)
    plan = HIPMolData(m, beam_table(make_dc_mot_beams(m, -0.5, 0.4, 2.0)), 1.0, hbar=HBAR)
    rng = np.random.default_rng(seed)
    n = 700
    z0 = torch.from_numpy(rng.normal(0, 1e-3, n)).to(dev)
    v0 = torch.from_numpy(rng.uniform(-15, 15, n)).to(dev)
    print(f"queue sizes: {ng}+{ne}", flush=True)
    g = traj_euler_device(plan, z0, v0, 0.2, DT, 1200, 0.015, kick_scale=1e-3, seed=4, schedule="grid")
    q = traj_euler_device(plan, z0, v0, 0.2, DT, 1200, 0.015, kick_scale=1e-3, seed=4, schedule="queue")
    torch.cuda.synchronize()
    assert (q[2].cpu().numpy() >= 0).all()
    assert _same(g, q)


def test_queue_dense_fallback(mol, dev):
    """A single sigma+ beam leaves ground states dark, so every step takes
    the dense elimination (the fallback call of the step loop): queue = one
    thread per particle bit for bit, and the oracle's trajectories."""
    import torch
    from molmot_amd.hip import HIPMolData, traj_euler_device
    from molmot_amd.obe.fields import LaserBeam
    bt = beam_table([LaserBeam(1.0, mol.omega_J32 - 3e6, 2, 3.0)])
    plan = HIPMolData(mol, bt, 1.0, hbar=HBAR)
    rng = np.random.default_rng(9)
    n = 150
    z0h, v0h = rng.normal(0, 1e-3, n), rng.uniform(-3, 3, n)
    z0, v0 = torch.from_numpy(z0h).to(dev), torch.from_numpy(v0h).to(dev)
    g = traj_euler_device(plan, z0, v0, 0.16, DT, 600, 0.015, schedule="grid")
    q = traj_euler_device(plan, z0, v0, 0.16, DT, 600, 0.015, schedule="queue")
    assert _same(g, q)
    zo, vo, nao, _, _ = oracle.traj_euler(mol, bt, z0h, v0h, DT, 600, 0.16, hbar=HBAR)
    zf, vf, na = (t.cpu().numpy() for t in q[:3])
    assert np.array_equal(na, nao)
    assert np.max(np.abs(zf - zo)) < 1e-10 and np.max(np.abs(vf - vo)) < 1e-8
