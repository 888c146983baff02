"""Gradient-driven parameter search against the reference's grid scan.

optimize_parameters (molmot/mot/force_scan.py:191-301) keeps the best spring
constant of a (detuning, split) grid.  The gradient search over the same box
(9 starts x 40 Adam steps, each step one forward + adjoint of two force
points) must reach at least the best k of a 21 x 21 grid of the reference's
rule (to 1%); freeing the four frequency components' detunings separately
must do no worse than the tied search."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def test_gradient_search_matches_or_beats_grid(mol, dev):
    from molmot_amd.hip.rate_equations import optimize_parameters_hip
    from molmot_amd.optimize import optimize_parameters_grad
    grid = optimize_parameters_hip(mol, np.linspace(-2.0, 0.0, 21), np.linspace(0.0, 2.0, 21), 1.0, 0.16)
    res = optimize_parameters_grad(mol, 0.16, (-2.0, 0.0), (0.0, 2.0), steps=40)
    print("grid", grid["k_best"], grid["delta_best"], grid["split_best"], flush=True)
    print("grad", res["k_best"], res["delta_best"], res["split_best"], res["evaluations"], flush=True)
    assert -2.0 - 1e-9 <= res["delta_best"] <= 1e-9 and -1e-9 <= res["split_best"] <= 2.0 + 1e-9
    assert res["feasible"] and res["beta_best"] > 0
    assert res["k_best"] >= 0.99 * grid["k_best"]
    ind = optimize_parameters_grad(mol, 0.16, (-2.0, 0.0), (0.0, 2.0), steps=40, independent=True,
                                   starts=[(res["delta_best"] + 1.0, res["split_best"] - 1.0)])
    print("independent", ind["k_best"], ind["theta"][:5], flush=True)
    assert ind["k_best"] >= 0.999 * res["k_best"]


def test_gradient_search_keeps_positive_damping(mol, dev):
    """On delta in [-3, -0.5] Gamma, split in [0, 2] Gamma the largest k of
    the grid sits where beta < 0 (delta = -0.5, split ~ 1.4); the reference's
    rule (force_scan.py:256-262) skips it, and so must the gradient search:
    a feasible optimum with k > 0, beta > 0 and k at least the grid's best
    admissible k (to 1%)."""
    from molmot_amd.hip.rate_equations import optimize_parameters_hip
    from molmot_amd.optimize import optimize_parameters_grad
    grid = optimize_parameters_hip(mol, np.linspace(-3.0, -0.5, 21), np.linspace(0.0, 2.0, 21), 1.0, 0.16)
    K, beta = grid["K_map"], grid["beta_map"]
    ib = np.unravel_index(np.argmax(K), K.shape)
    print("grid max k", K[ib], "beta there", beta[ib], "admissible best", grid["k_best"], grid["beta_best"],
          flush=True)
    assert beta[ib] < 0 and K[ib] > grid["k_best"] > 0 and grid["beta_best"] > 0
    res = optimize_parameters_grad(mol, 0.16, (-3.0, -0.5), (0.0, 2.0), steps=40)
    print("grad", res["k_best"], res["beta_best"], res["delta_best"], res["split_best"], flush=True)
    assert res["feasible"] and res["k_best"] > 0 and res["beta_best"] > 0
    assert res["k_best"] >= 0.99 * grid["k_best"]
    assert -3.0 - 1e-9 <= res["delta_best"] <= -0.5 + 1e-9
