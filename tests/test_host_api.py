"""Host-side drop-in surfaces (CPU only): problem set-up, beam pairs, the 3D
beam list, and agreement of the product's table construction with the
oracle's restatement."""
import numpy as np

from oracle import mcwf as mc
from molmot_amd.obe.beam_pairs import make_beam_pairs_dc4, make_beam_pairs_rf, make_beam_pairs_blue


def test_sse_problem_matches_reference_shape(mol):
    """tests/test_all_functions.py:422-430 of the reference."""
    from molmot_amd.obe.stochastic import SSEProblem
    bp = make_beam_pairs_dc4(mol, -0.30, 0.40, 1.0)
    prob = SSEProblem(mol, bp, B_gradient=0.16)
    assert prob.n_states == 16 and prob.state_size == 2 * 16 + 6
    u0 = prob.build_initial_state([1e-3, 0, 0], [0, 0, 1.0])
    assert u0[0] == 1.0 and abs(u0[prob.r_idx] - 1e-3 * mol.k) < 1e-9 and abs(u0[prob.v_idx + 2] * prob.v_nat - 1) < 1e-15


def test_sse_tables_equal_oracle(mol):
    from molmot_amd.obe.stochastic import SSEProblem, polarization_table
    bp = make_beam_pairs_dc4(mol, -0.5, 0.6, 2.0)
    prob = SSEProblem(mol, bp, 0.2, diffusion_constants=[1e-50, 2e-50, 3e-50])
    tab = mc.tables(mol, bp, 0.2, diffusion_constants=[1e-50, 2e-50, 3e-50])
    assert np.array_equal(polarization_table(prob.pols), tab["eps"])
    assert np.array_equal(prob.d_ge, tab["d_ge"]) and np.allclose(prob.branching, tab["branching"], rtol=1e-15)
    assert np.array_equal(prob.omega_lasers, tab["omega_lasers"])
    assert np.allclose(prob.diffusion_nat, tab["diffusion_nat"], rtol=1e-15)
    assert abs(prob.mass_nat / tab["mass_nat"] - 1) < 1e-15
    np.testing.assert_allclose(prob.get_omega0s_zeeman(np.array([300.0, -200.0, 100.0])),
                               mc.omega0s_zeeman(tab, np.array([300.0, -200.0, 100.0])), rtol=1e-15)


def test_beam_pairs(mol):
    G = mol.Gamma
    dc = make_beam_pairs_dc4(mol, -0.3, 0.4, 1.5)
    assert [b["q_fwd"] for b in dc] == [2, 0, 2, 0] and [b["q_bwd"] for b in dc] == [0, 2, 0, 2]
    assert dc[0]["freq"] == mol.omega_J32 + (-0.3 * G / (2 * np.pi)) + 0.4 * G / (2 * np.pi)
    rf = make_beam_pairs_rf(mol, -1.0, 2.0, phase=1)
    assert rf[0]["q_fwd"] == 0 and rf[1]["q_fwd"] == 2 and rf[0]["s0"] == 2.0
    bl = make_beam_pairs_blue(mol, 1.0, 0.5, 1.0)
    assert bl[1]["freq"] == mol.omega_J12 + G / (2 * np.pi) + 0.5 * G / (2 * np.pi)


def test_six_beam_config(mol):
    from molmot_amd.mot.simulator_3d import MOTSimulator3D
    from oracle import sim3d as s3
    bp = make_beam_pairs_dc4(mol, -0.30, 0.40, 1.0)
    sim = MOTSimulator3D(mol, bp, B_gradient=0.16)
    ob = s3.six_beams(bp)
    assert len(sim.beams) == 24 == len(ob)
    for b, o in zip(sim.beams, ob):
        assert np.array_equal(b.k_hat, o["k_hat"]) and b.q_lab == o["q"] and b.freq == o["freq"]


def test_gradient_search_keeps_the_reference_rule():
    """gradient_search with (ascent, merit, kind): the unconstrained maximum
    of k(u) = -(u - 0.8)^2 lies where the damping alpha(u) = 0.3 - u is
    negative; with the reference's rule (merit = k only where k > 0 ... here
    k + 1 > 0 and alpha > 0) the best point is just inside alpha > 0."""
    import torch
    from molmot_amd.optimize import gradient_search

    def obj(th):
        u = th[0]
        k = 1.0 - (u - 0.8) ** 2
        a = 0.3 - u
        kv, av = float(k), float(a)
        merit = kv if (kv > 0 and av > 0) else 0.0
        return (k, merit, "k") if av > 0 else (a, merit, "alpha")

    D = np.zeros((1, 7))
    D[0, 0] = 1.0
    r = gradient_search(np.zeros(7), obj, D, [1.0], lower=[-1.0], upper=[1.0], steps=200, lr=0.02)
    u = r["theta"][0]
    assert u < 0.3 and u > 0.25
    assert abs(r["merit"] - (1.0 - (u - 0.8) ** 2)) < 1e-15
    # plain objective: the unconstrained maximum
    r2 = gradient_search(np.zeros(7), lambda th: 1.0 - (th[0] - 0.8) ** 2, D, [1.0], lower=[-1.0], upper=[1.0],
                         steps=300, lr=0.02)
    assert abs(r2["theta"][0] - 0.8) < 1e-2
