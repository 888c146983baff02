"""GPU adjoint kernel (mm_rate_eq_grad_points) and the autograd Function
against the torch-autograd FP64 oracle (itself pinned to finite differences
of the C oracle in tests/test_oracle_gradient.py).  Bar: 1e-8 norm-relative
per Jacobian column (the north-star bar for gradients is 1e-6)."""
import numpy as np
import pytest

import oracle
from oracle import torch_rate_eq
from molmot_amd import configs
from molmot_amd.constants import hbar as HBAR
from molmot_amd.obe.fields import beam_table, make_dc_mot_beams

pytestmark = pytest.mark.gpu

B0 = configs.B_GRAD_SI


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def colrel(a, b):
    return np.max(np.abs(a - b), axis=0) / np.maximum(np.max(np.abs(b), axis=0), 1e-300)


@pytest.mark.parametrize("cfg", ["red_mot", "dc"])
def test_jacobian_vs_torch_oracle(mol, dev, cfg):
    import torch
    from molmot_amd.adjoint import ForceProblem, force_and_jacobian
    if cfg == "red_mot":
        prob, B = ForceProblem(mol, configs.red_mot_beams(mol)), B0
    else:
        prob, B = ForceProblem(mol, make_dc_mot_beams(mol, -0.3, 0.4, 1.0)), 0.16
    rng = np.random.default_rng(7)
    v = rng.uniform(-25, 25, 200)
    z = rng.uniform(-9e-3, 9e-3, 200)
    theta = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 1.0, B])
    F, J = force_and_jacobian(prob, theta, torch.from_numpy(v).to(dev), torch.from_numpy(z).to(dev))
    Jt, Ft = torch_rate_eq.jacobian(mol, beam_table(prob.beams), theta, v, z, hbar=HBAR)
    assert np.max(np.abs(F.cpu().numpy() - Ft)) / np.max(np.abs(Ft)) < 1e-12
    err = colrel(J.cpu().numpy(), Jt)
    assert np.all(err < 1e-8), err


def test_jacobian_full_map_consistency(mol, dev):
    """1024 x 1024 map (v in [-30,30] m/s, z in [-10,10] mm): the adjoint
    kernel's force equals the forward map kernel's, every gradient is finite,
    and dF/dz, dF/dv agree with the forward map's own central differences."""
    import torch
    from molmot_amd.adjoint import ForceProblem, force_and_jacobian
    from molmot_amd.hip import rate_eq_map_device
    prob = ForceProblem(mol, configs.red_mot_beams(mol))
    vg, zg = configs.force_map_grid()
    V, Z = np.meshgrid(vg, zg, indexing="ij")
    theta = prob.theta0(B0).numpy()
    F, J = force_and_jacobian(prob, theta, torch.from_numpy(V.ravel()).to(dev), torch.from_numpy(Z.ravel()).to(dev))
    Fm, _, _ = rate_eq_map_device(prob.plan(theta), torch.from_numpy(vg).to(dev), torch.from_numpy(zg).to(dev), B0,
                                  want_pop=False, want_scatter=False)
    assert torch.max(torch.abs(F - Fm)) / torch.max(torch.abs(Fm)) < 1e-13
    assert bool(torch.isfinite(J).all())
    # spot check 512 points against the oracle Jacobian
    rng = np.random.default_rng(3)
    idx = rng.integers(0, V.size, 512)
    Jt, _ = torch_rate_eq.jacobian(mol, beam_table(prob.beams), theta, V.ravel()[idx], Z.ravel()[idx], hbar=HBAR)
    assert np.all(colrel(J.cpu().numpy()[idx], Jt) < 1e-8)


def test_autograd_spring_damping_vs_fd(mol, dev):
    """d k / d theta and d alpha / d theta through torch.autograd (backward =
    adjoint kernel) vs central differences of the GPU forward at exactly
    representable steps."""
    import torch
    from molmot_amd.adjoint import ForceProblem, spring_constant_t, damping_coefficient_t
    prob = ForceProblem(mol, make_dc_mot_beams(mol, -0.3, 0.4, 1.0))
    theta = prob.theta0(0.16).requires_grad_(True)
    k = spring_constant_t(theta, prob)
    (gk,) = torch.autograd.grad(k, theta)
    a = damping_coefficient_t(theta, prob)
    (ga,) = torch.autograd.grad(a, theta)
    steps = [64.0] * 5 + [1e-5]
    for j, h in enumerate(steps):
        tp = theta.detach().clone()
        tm = theta.detach().clone()
        tp[j] += h
        tm[j] -= h
        fdk = (spring_constant_t(tp, prob) - spring_constant_t(tm, prob)).item() / (2 * h)
        fda = (damping_coefficient_t(tp, prob) - damping_coefficient_t(tm, prob)).item() / (2 * h)
        assert abs(gk[j].item() - fdk) <= 1e-5 * max(abs(fdk), 1e-30) + 1e-12 * abs(k.item()), (j, gk[j].item(), fdk)
        assert abs(ga[j].item() - fda) <= 1e-5 * max(abs(fda), 1e-30) + 1e-12 * abs(a.item()), (j, ga[j].item(), fda)


def test_autograd_force_backward_shapes(mol, dev):
    import torch
    from molmot_amd.adjoint import ForceProblem, rate_force
    prob = ForceProblem(mol, configs.red_mot_beams(mol))
    theta = prob.theta0(B0).requires_grad_(True)
    v = torch.linspace(-5, 5, 33, dtype=torch.float64, device=dev).requires_grad_(True)
    z = torch.linspace(-3e-3, 3e-3, 33, dtype=torch.float64, device=dev).requires_grad_(True)
    F = rate_force(theta, v, z, prob)
    (F * F).sum().backward()
    assert theta.grad.shape == (7,) and v.grad.shape == (33,) and z.grad.shape == (33,)
    assert bool(torch.isfinite(theta.grad).all())


def test_dark_points_give_nan_gradient(mol, dev):
    import torch
    from molmot_amd.adjoint import ForceProblem, force_and_jacobian
    prob = ForceProblem(mol, make_dc_mot_beams(mol, -0.2, 0.7, 0.0))
    F, J = force_and_jacobian(prob, np.array([0, 0, 0, 0, 0, 1.0, 0.16]),
                              torch.zeros(3, dtype=torch.float64, device=dev),
                              torch.zeros(3, dtype=torch.float64, device=dev))
    assert bool(torch.isnan(J).all()) and bool((F == 0).all())


def test_jacobian_forty_levels_vs_torch_oracle(dev):
    """Seeded synthetic 36+4 molecule (the 40-level class), DC beams."""
    import torch
    from molmot_amd.adjoint import ForceProblem, force_and_jacobian
    from molmot_amd.molecules import synthetic_molecule
    m = synthetic_molecule(36, 4, seed=5)
    prob = ForceProblem(m, make_dc_mot_beams(m, -0.5, 0.6, 2.0))
    rng = np.random.default_rng(3)
    v = rng.uniform(-10, 10, 64)
    z = rng.uniform(-6e-3, 6e-3, 64)
    theta = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.16])
    F, J = force_and_jacobian(prob, theta, torch.from_numpy(v).to(dev), torch.from_numpy(z).to(dev))
    Jt, Ft = torch_rate_eq.jacobian(m, beam_table(prob.beams), theta, v, z, hbar=HBAR)
    Fg = F.cpu().numpy()
    assert np.all(np.isfinite(Fg))
    assert np.max(np.abs(Fg - Ft)) / np.max(np.abs(Ft)) < 1e-12
    err = colrel(J.cpu().numpy(), Jt)
    assert np.all(err < 1e-8), err


def test_jacobian_general_term_path(mol, dev):
    """The red-MOT beams with their four frequency components relabelled as
    two of equal split sign (c -> c % 2; the split signs are +, -, +, -):
    every pair then holds two terms of each component, so the plan has no
    component-ordered records and the gradient kernel walks the general term
    list.  It must agree with the torch oracle for the relabelled beams, give
    the same force as the component-ordered kernel on the original beams, and
    its merged detuning columns must be the sums of the original ones (one
    offset moves both components)."""
    import dataclasses
    import torch
    from molmot_amd.adjoint import ForceProblem, force_and_jacobian
    beams = configs.red_mot_beams(mol)
    bt = beam_table(beams)
    for c in range(4):
        assert np.all(bt.split_sign[bt.comp == c] == bt.split_sign[bt.comp == c % 2][0])
    merged = [dataclasses.replace(b, comp=int(b.comp) % 2) for b in beams]
    assert beam_table(merged).n_comp == 2
    rng = np.random.default_rng(11)
    v = rng.uniform(-20, 20, 256)
    z = rng.uniform(-8e-3, 8e-3, 256)
    theta = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 1.0, B0])
    vt, zt = torch.from_numpy(v).to(dev), torch.from_numpy(z).to(dev)
    F1, J1 = force_and_jacobian(ForceProblem(mol, beams), theta, vt, zt)
    F2, J2 = force_and_jacobian(ForceProblem(mol, merged), theta, vt, zt)
    F1, J1, F2, J2 = (x.cpu().numpy() for x in (F1, J1, F2, J2))
    assert np.max(np.abs(F2 - F1)) <= 1e-14 * np.max(np.abs(F1))
    Jt, _ = torch_rate_eq.jacobian(mol, beam_table(merged), theta, v, z, hbar=HBAR)
    err = colrel(J2[:, [0, 1, 4, 5, 6, 7, 8]], Jt[:, [0, 1, 4, 5, 6, 7, 8]])
    assert np.all(err < 1e-8), err
    assert np.all(J2[:, 2:4] == 0.0)
    Jsum = np.stack([J1[:, 0] + J1[:, 2], J1[:, 1] + J1[:, 3]], axis=1)
    assert np.all(colrel(J2[:, 0:2], Jsum) < 1e-10)
    assert np.all(colrel(J2[:, 4:], J1[:, 4:]) < 1e-10)
