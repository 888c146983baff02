"""HIP-graph capture of C-ABI launch sequences.

Each replay runs with new inputs written in place, with eager GPU work and a
host synchronisation between replays, and must reproduce the eager result of
the same launches bit for bit: the checkpointed trajectory adjoint (forward
kernel, one kernel per segment, two strided copies) and a force map followed
by a trajectory batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


THETA = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.16])


def test_adjoint_graph_replays_match_eager(mol, dev):
    import torch
    from molmot_amd.adjoint import ForceProblem, trajectory_vjp
    from molmot_amd.hip.graphs import TrajectoryAdjointGraph
    from molmot_amd.hip.rate_equations import rate_eq_map_device
    from molmot_amd.obe.fields import make_dc_mot_beams
    prob = ForceProblem(mol, make_dc_mot_beams(mol, -0.20, 0.70, 1.0))
    n, nu, K = 96, 300, 64
    g = TrajectoryAdjointGraph(prob, THETA, n, mol.mass, 1e-5, nu, 8e-3, K, dev)
    rng = np.random.default_rng(5)
    sets = []
    for _ in range(3):
        sets.append([torch.tensor(x, device=dev) for x in (rng.uniform(-6e-3, 6e-3, n), rng.uniform(-8, 8, n),
                                                          rng.normal(size=n), rng.normal(size=n) * 1e-3)])
    plan = prob.plan(THETA)
    vv = torch.linspace(-2, 2, 64, dtype=torch.float64, device=dev)
    for rep in (0, 1, 2, 0):
        z0, v0, cz, cv = sets[rep]
        for dst, src in zip((g.z0, g.v0, g.cot_z, g.cot_v), (z0, v0, cz, cv)):
            dst.copy_(src)
        g.replay()
        torch.cuda.synchronize()
        ref = trajectory_vjp(prob, THETA, z0, v0, mol.mass, 1e-5, nu, cz, cv, 8e-3, None, K)
        torch.cuda.synchronize()
        assert torch.equal(g.z_final, ref["z_final"]) and torch.equal(g.v_final, ref["v_final"])
        assert torch.equal(g.n_done, ref["n_done"])
        assert torch.equal(g.g_theta, ref["g_theta"])
        assert torch.equal(g.g_z0, ref["g_z0"]) and torch.equal(g.g_v0, ref["g_v0"])
        # eager work between replays
        rate_eq_map_device(plan, vv, vv * 1e-3, 0.16)
        torch.cuda.synchronize()


def test_captured_map_then_trajectories(mol, dev):
    import torch
    from molmot_amd.hip import _lib
    from molmot_amd.hip.graphs import CapturedLaunches
    from molmot_amd.hip.plan import HIPMolData
    from molmot_amd.hip.rate_equations import rate_eq_map_device
    from molmot_amd.hip.trajectories import traj_euler_device
    from molmot_amd.obe.fields import make_dc_mot_beams
    plan = HIPMolData(mol, make_dc_mot_beams(mol, -0.2, 0.7, 1.0))
    f64 = dict(dtype=torch.float64, device=dev)
    v, z = torch.zeros(48, **f64), torch.zeros(40, **f64)
    out = (torch.empty(48 * 40, **f64), torch.empty((48 * 40, 12), **f64), torch.empty(48 * 40, **f64))
    n = 128
    z0, v0 = torch.zeros(n, **f64), torch.zeros(n, **f64)
    zf, vf = torch.empty(n, **f64), torch.empty(n, **f64)
    na = torch.empty(n, dtype=torch.int32, device=dev)

    def launches():
        rate_eq_map_device(plan, v, z, 0.16, out=out)
        rc = _lib.lib().mm_traj_euler_1d(plan.handle, n, _lib.ptr(z0), _lib.ptr(v0), 0.16, plan.mass, 1e-5, 400,
                                         0.015, None, None, _lib.ptr(zf), _lib.ptr(vf), _lib.ptr(na), 1e-3, 42, 0, 7,
                                         _lib.stream_handle())
        _lib.check(rc, "mm_traj_euler_1d")

    cap = CapturedLaunches(launches)
    rng = np.random.default_rng(9)
    for _ in range(3):
        v.copy_(torch.tensor(rng.uniform(-3, 3, 48), device=dev))
        z.copy_(torch.tensor(rng.uniform(-5e-3, 5e-3, 40), device=dev))
        z0.copy_(torch.tensor(rng.uniform(-5e-3, 5e-3, n), device=dev))
        v0.copy_(torch.tensor(rng.uniform(-5, 5, n), device=dev))
        cap.replay()
        torch.cuda.synchronize()
        F, P, R = rate_eq_map_device(plan, v, z, 0.16)
        zr, vr, nr, _, _ = traj_euler_device(plan, z0, v0, 0.16, 1e-5, 400, 0.015, kick_scale=1e-3, seed=42,
                                             stream_id=7)
        torch.cuda.synchronize()
        assert torch.equal(out[0], F) and torch.equal(out[1], P) and torch.equal(out[2], R)
        assert torch.equal(zf, zr) and torch.equal(vf, vr) and torch.equal(na, nr)
