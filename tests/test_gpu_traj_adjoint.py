"""Checkpointed trajectory adjoint (mm_traj_vjp_1d) and the capture-velocity
gradient.

* VJP through 300 Euler steps of 4 particles (one escaping) equals torch FP64
  autograd through the CPU restatement of the force (oracle.torch_rate_eq):
  1e-8 norm-relative on g_theta, g_z0, g_v0.
* Independent of the checkpoint interval K (1, 7, 64, 1000): 1e-12.
* Forward sensitivities (mm_traj_tangent_1d) contracted with the same
  cotangents equal the torch oracle (1e-8) and the checkpointed adjoint
  (1e-10, both directions of the 2 x 9 Jacobian).
* Autograd Function (trajectory_final) in both modes matches the raw VJP.
* dv*/dtheta of the continuous capture threshold matches central finite
  differences of the threshold itself (5e-3 relative per component: the
  discrete turning step switches between neighbouring states across the
  finite-difference interval)."""
import numpy as np
import pytest

from oracle import torch_rate_eq
from molmot_amd.constants import hbar as HBAR
from molmot_amd.obe.fields import beam_table, make_dc_mot_beams

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def _problem(mol):
    from molmot_amd.adjoint import ForceProblem
    return ForceProblem(mol, make_dc_mot_beams(mol, -0.20, 0.70, 1.0))


Z0 = np.array([3e-3, -1e-3, 0.5e-3, 5e-3])
V0 = np.array([-4.0, 2.0, 0.3, 12.0])
CZ = np.array([1.0, -0.5, 2.0, 0.7])
CV = np.array([1e-4, 3e-4, -2e-4, 1e-4])
THETA = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.16])
DT, NU, ZESC = 1e-5, 300, 8e-3


def _torch_oracle(mol, prob):
    import torch
    th = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
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
    L = (torch.tensor(CZ) * z + torch.tensor(CV) * v).sum()
    L.backward()
    return th.grad.numpy(), z0.grad.numpy(), v0.grad.numpy(), z.detach().numpy(), v.detach().numpy(), done


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_vjp_vs_torch_oracle(mol, dev):
    import torch
    from molmot_amd.adjoint import trajectory_vjp
    prob = _problem(mol)
    o = trajectory_vjp(prob, THETA, torch.tensor(Z0, device=dev), torch.tensor(V0, device=dev), mol.mass, DT, NU,
                       torch.tensor(CZ, device=dev), torch.tensor(CV, device=dev), ZESC, None, 64)
    gth, gz0, gv0, zf, vf, done = _torch_oracle(mol, prob)
    nd = o["n_done"].cpu().numpy()
    assert np.array_equal(nd, done) and nd[3] < NU and nd[1] == NU
    assert _rel(o["z_final"].cpu().numpy(), zf) < 1e-12 and _rel(o["v_final"].cpu().numpy(), vf) < 1e-12
    g = o["g_theta"].sum(0).cpu().numpy()
    for j in range(7):
        assert abs(g[j] - gth[j]) <= 1e-8 * max(abs(gth[j]), 1e-300), (j, g[j], gth[j])
    assert _rel(o["g_z0"].cpu().numpy(), gz0) < 1e-8
    assert _rel(o["g_v0"].cpu().numpy(), gv0) < 1e-8


def test_vjp_checkpoint_interval_independent(mol, dev):
    import torch
    from molmot_amd.adjoint import trajectory_vjp
    prob = _problem(mol)
    outs = []
    for K in (1, 7, 64, 1000):
        o = trajectory_vjp(prob, THETA, torch.tensor(Z0, device=dev), torch.tensor(V0, device=dev), mol.mass, DT,
                           NU, torch.tensor(CZ, device=dev), torch.tensor(CV, device=dev), ZESC, None, K)
        outs.append(np.concatenate([o["g_theta"].cpu().numpy().ravel(), o["g_z0"].cpu().numpy(),
                                    o["g_v0"].cpu().numpy()]))
    for x in outs[1:]:
        assert _rel(x, outs[0]) < 1e-12


def test_tangent_vs_torch_oracle_and_checkpoint(mol, dev):
    import torch
    from molmot_amd.adjoint import trajectory_tangent, trajectory_vjp, vjp_from_sens
    prob = _problem(mol)
    z0, v0 = torch.tensor(Z0, device=dev), torch.tensor(V0, device=dev)
    o = trajectory_tangent(prob, THETA, z0, v0, mol.mass, DT, NU, ZESC)
    gth, gz0, gv0, zf, vf, done = _torch_oracle(mol, prob)
    assert np.array_equal(o["n_done"].cpu().numpy(), done)
    assert _rel(o["z_final"].cpu().numpy(), zf) < 1e-12 and _rel(o["v_final"].cpu().numpy(), vf) < 1e-12
    g, g_z0, g_v0 = vjp_from_sens(o["sens"], torch.tensor(CZ, device=dev), torch.tensor(CV, device=dev))
    g = g.sum(0).cpu().numpy()
    for j in range(7):
        assert abs(g[j] - gth[j]) <= 1e-8 * max(abs(gth[j]), 1e-300), (j, g[j], gth[j])
    assert _rel(g_z0.cpu().numpy(), gz0) < 1e-8 and _rel(g_v0.cpu().numpy(), gv0) < 1e-8
    # each row of the Jacobian against the checkpointed adjoint with a unit cotangent
    one, zero = torch.ones(4, dtype=torch.float64, device=dev), torch.zeros(4, dtype=torch.float64, device=dev)
    for row, (cz, cv) in enumerate(((one, zero), (zero, one))):
        a = trajectory_vjp(prob, THETA, z0, v0, mol.mass, DT, NU, cz, cv, ZESC, None, 64)
        ref = torch.cat([a["g_theta"], a["g_z0"][:, None], a["g_v0"][:, None]], 1).cpu().numpy()
        got = o["sens"][:, row].cpu().numpy()
        for i in range(4):
            assert _rel(got[i], ref[i]) < 1e-10, (row, i)


def test_autograd_function(mol, dev):
    import torch
    from molmot_amd.adjoint import trajectory_final, trajectory_vjp
    prob = _problem(mol)
    o = trajectory_vjp(prob, THETA, torch.tensor(Z0[:3], device=dev), torch.tensor(V0[:3], device=dev), mol.mass,
                       DT, NU, torch.tensor(CZ[:3], device=dev), torch.tensor(CV[:3], device=dev), ZESC, None, 50)
    for mode, tol in (("checkpoint", 1e-14), ("tangent", 1e-10)):
        th = torch.tensor(THETA, requires_grad=True)
        z0 = torch.tensor(Z0[:3], device=dev, requires_grad=True)
        v0 = torch.tensor(V0[:3], device=dev, requires_grad=True)
        zf, vf = trajectory_final(th, z0, v0, prob, mol.mass, DT, NU, ZESC, 50, mode=mode)
        (zf * torch.tensor(CZ[:3], device=dev) + vf * torch.tensor(CV[:3], device=dev)).sum().backward()
        assert _rel(th.grad.numpy(), o["g_theta"].sum(0).cpu().numpy()) < tol, mode
        assert _rel(v0.grad.cpu().numpy(), o["g_v0"].cpu().numpy()) < tol, mode
        assert _rel(z0.grad.cpu().numpy(), o["g_z0"].cpu().numpy()) < tol, mode


def test_capture_velocity_gradient_vs_fd(mol, dev):
    from molmot_amd.adjoint import capture_threshold, capture_velocity_grad
    prob = _problem(mol)
    kw = dict(z_start=3e-3, v_max=30.0, dt=2e-6, n_updates=6000, z_escape=0.015)
    vstar, g = capture_velocity_grad(prob, THETA, ckpt_every=500, **kw)
    print("v* =", vstar, "grad =", g, flush=True)
    v2, g2 = capture_velocity_grad(prob, THETA, mode="checkpoint", **kw)
    assert v2 == vstar and _rel(g2, g) < 1e-9
    assert 0.5 < vstar < 30.0
    steps = [2e4, 2e4, 2e4, 2e4, 2e4, 2e-3, 2e-4]
    for j, h in enumerate(steps):
        tp, tm = THETA.copy(), THETA.copy()
        tp[j] += h
        tm[j] -= h
        fd = (capture_threshold(prob, tp, tol=1e-10, **kw) - capture_threshold(prob, tm, tol=1e-10, **kw)) / (2 * h)
        print(j, g[j], fd, flush=True)
        assert abs(g[j] - fd) <= 5e-3 * max(abs(fd), abs(g[j])) + 1e-9 / h, (j, g[j], fd)


def test_prebuilt_plan_argument(mol, dev):
    """trajectory_tangent, trajectory_vjp and force_and_jacobian with a plan
    the caller built (problem.plan(theta), reused across calls) return what
    the calls building their own plan return, bit for bit."""
    import torch
    from molmot_amd.adjoint import force_and_jacobian, trajectory_tangent, trajectory_vjp
    prob = _problem(mol)
    B0 = 0.16
    theta = prob.theta0(B0).numpy()
    plan = prob.plan(theta)
    z0, v0 = torch.from_numpy(Z0).to(dev), torch.from_numpy(V0).to(dev)
    cz = torch.from_numpy(CZ).to(dev)
    a = trajectory_tangent(prob, theta, z0, v0, mol.mass, 1e-5, 300, 0.015, None)
    b = trajectory_tangent(prob, theta, z0, v0, mol.mass, 1e-5, 300, 0.015, None, plan=plan)
    assert all(torch.equal(a[k], b[k]) for k in a)
    a = trajectory_vjp(prob, theta, z0, v0, mol.mass, 1e-5, 300, cz, None, 0.015, None, 64)
    b = trajectory_vjp(prob, theta, z0, v0, mol.mass, 1e-5, 300, cz, None, 0.015, None, 64, plan=plan)
    assert all(torch.equal(a[k], b[k]) for k in a)
    v = torch.linspace(-3, 3, 50, dtype=torch.float64, device=dev)
    z = torch.linspace(-2e-3, 2e-3, 50, dtype=torch.float64, device=dev)
    fa, ja = force_and_jacobian(prob, theta, v, z)
    fb, jb = force_and_jacobian(prob, theta, v, z, plan=plan)
    assert torch.equal(fa, fb) and torch.equal(ja.nan_to_num(7.0), jb.nan_to_num(7.0))
