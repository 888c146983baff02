"""Pin the torch-autograd gradient oracle (CPU): its force equals the C
oracle's, and its Jacobian equals central finite differences of the C oracle
(steps ~1e-6 relative to each parameter's scale, frequency steps chosen
exactly representable)."""
import numpy as np
import pytest
import torch

import oracle
from oracle import torch_rate_eq
from molmot_amd.adjoint import ForceProblem
from molmot_amd.constants import hbar as HBAR
from molmot_amd import configs
from molmot_amd.obe.fields import beam_table

B0 = configs.B_GRAD_SI


@pytest.fixture(scope="module")
def problem(mol):
    return ForceProblem(mol, configs.red_mot_beams(mol))


def c_force(problem, theta, v, z):
    bt = problem.table(theta)
    F, _, _ = oracle.rate_eq_points(problem.mol_data, bt, v, z, float(theta[6]), hbar=HBAR, flavour="jit")
    return F


def test_torch_force_matches_c_oracle(mol, problem):
    rng = np.random.default_rng(0)
    v = rng.uniform(-20, 20, 64)
    z = rng.uniform(-8e-3, 8e-3, 64)
    theta = np.array([1e5, -2e5, 3e4, 0.0, 5e4, 1.1, B0])
    Ft = torch_rate_eq.force(mol, beam_table(problem.beams), torch.as_tensor(theta), torch.as_tensor(v),
                             torch.as_tensor(z)).numpy()
    Fc = c_force(problem, theta, v, z)
    assert np.max(np.abs(Ft - Fc)) / np.max(np.abs(Fc)) < 1e-13


def test_torch_jacobian_matches_central_fd(mol, problem):
    rng = np.random.default_rng(1)
    v = rng.uniform(-20, 20, 24)
    z = rng.uniform(-8e-3, 8e-3, 24)
    theta0 = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 1.0, B0])
    J, F = torch_rate_eq.jacobian(mol, beam_table(problem.beams), theta0, v, z, hbar=HBAR)
    # steps ~1e-6 of each parameter's scale; frequency steps are 8 Hz = 128 ulp
    # of the ~4.3e14 Hz beam frequencies so the perturbed frequencies are
    # exact (a 6.4 Hz step would be realised as 6.375 Hz)
    steps = np.array([8.0] * 5 + [1e-6])
    for j in range(6):
        h = steps[j]
        tp = theta0.copy()
        tm = theta0.copy()
        tp[j] += h
        tm[j] -= h
        fd = (c_force(problem, tp, v, z) - c_force(problem, tm, v, z)) / (2 * h)
        err = np.max(np.abs(J[:, j] - fd)) / np.max(np.abs(fd))
        assert err < 1e-6, (j, err)


def test_torch_jacobian_field_and_kinematic_directions(mol, problem):
    """B', z and v enter the detuning through the Zeeman and Doppler shifts
    of a ~4.3e14 Hz transition whose energies carry 0.0625 Hz ulps, so in
    the reference's own arithmetic F is a staircase in those variables at
    that resolution: a 1e-6 relative step moves the shift by only a few Hz
    and the finite difference is dominated by the steps.  These three
    directions are checked with Richardson-extrapolated central differences,
    (4 D(h/2) - D(h)) / 3, at h = 3e-3 of the scale, to 5e-6."""
    rng = np.random.default_rng(1)
    v = rng.uniform(-20, 20, 24)
    z = rng.uniform(-8e-3, 8e-3, 24)
    theta0 = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 1.0, B0])
    J, _ = torch_rate_eq.jacobian(mol, beam_table(problem.beams), theta0, v, z, hbar=HBAR)

    def d_B(h):
        tp, tm = theta0.copy(), theta0.copy()
        tp[6] += h
        tm[6] -= h
        return (c_force(problem, tp, v, z) - c_force(problem, tm, v, z)) / (2 * h)

    def d_z(h):
        return (c_force(problem, theta0, v, z + h) - c_force(problem, theta0, v, z - h)) / (2 * h)

    def d_v(h):
        return (c_force(problem, theta0, v + h, z) - c_force(problem, theta0, v - h, z)) / (2 * h)

    for col, fn, scale in [(6, d_B, B0), (7, d_z, 1e-2), (8, d_v, 1.0)]:
        h = 3e-3 * scale
        rich = (4 * fn(h / 2) - fn(h)) / 3
        err = np.max(np.abs(J[:, col] - rich)) / np.max(np.abs(J[:, col]))
        assert err < 5e-6, (col, err)
