"""GPU Monte Carlo wavefunction kernel vs the FP64 oracle.

* FP64 amplitudes: trajectories, populations, forces, jump steps and
  channels equal the oracle's replay of the same Philox stream (1e-9 relative
  on the state, identical jumps), with recoil, diffusion kicks and escape.
* FP32 amplitudes: the deterministic part (before the first jump) within
  2e-5 of the oracle.
* Ensemble: the mean normalised force of 16,384 FP32 trajectories moving at
  fixed velocity equals the density-matrix force of the master equation this
  MCWF unravels (oracle.mcwf.lindblad_path) within 5 standard errors plus
  2 % of the force scale, at v = 0, 0.5 and 1 m/s."""
import dataclasses

import numpy as np
import pytest

from oracle import mcwf as mc
from molmot_amd.obe.beam_pairs import make_beam_pairs_dc4

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def _problem(mol, s0=3.0, B=0.16, kick=True, diffusion=None, heavy=False):
    from molmot_amd.obe.stochastic import SSEProblem
    m = dataclasses.replace(mol, mass=mol.mass * 1e12) if heavy else mol
    bp = make_beam_pairs_dc4(m, -1.0, 0.5, s0)
    prob = SSEProblem(m, bp, B, add_spontaneous_kick=kick, diffusion_constants=diffusion)
    tab = mc.tables(m, bp, B, add_spontaneous_kick=kick, diffusion_constants=diffusion)
    return prob, tab


def _run_gpu(prob, r0, v0, psi0, n_steps, save_every, seed, precision, r_escape=0.0, jump_cap=0,
             jump_rule="reference"):
    import torch
    from molmot_amd.obe.stochastic import MCWFPlan, mcwf_ensemble_device
    plan = MCWFPlan(prob)
    d = torch.device("cuda", 0)
    ps = None if psi0 is None else torch.from_numpy(np.asarray(psi0, dtype=complex)).to(d)
    out = mcwf_ensemble_device(plan, torch.from_numpy(np.asarray(r0, float)).to(d),
                               torch.from_numpy(np.asarray(v0, float)).to(d), 0.01, n_steps, save_every, seed, ps,
                               precision, r_escape, 0, 0, jump_cap, jump_rule)
    torch.cuda.synchronize()
    return plan, {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()}


@pytest.mark.parametrize("rule", ["reference", "exact"])
def test_fp64_replays_oracle_with_jumps(mol, dev, rule):
    D = np.array([4e-50, 0.0, 1e-49])
    prob, tab = _problem(mol, diffusion=D)
    r0 = np.array([[0.0, 0.0, 0.0], [900.0, -400.0, 300.0], [-2000.0, 500.0, 1500.0], [200.0, 100.0, -50.0]])
    v0 = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, -0.05], [0.0, 0.2, 0.1], [-0.1, 0.1, 0.3]])
    psi0 = np.zeros((4, 16), dtype=complex)
    psi0[:, 4] = 1.0
    psi0[1] = 0.0
    psi0[1, [4, 7]] = [0.6, 0.8j]
    plan, g = _run_gpu(prob, r0, v0, psi0, 1500, 100, 1234, "fp64", jump_cap=64, jump_rule=rule)
    assert abs(plan.shift[0] - mc.default_shift(tab)[0]) == 0 and abs(plan.shift[1] - mc.default_shift(tab)[1]) == 0
    total_jumps = 0
    for i in range(4):
        print(f"trajectory {i}", flush=True)
        o = mc.run(tab, r0[i], v0[i], psi0[i], 0.01, 1500, 100, seed=1234, traj=i, shift=plan.shift,
                   jump_rule=int(rule == "exact"))
        assert g["n_scatters"][i] == o["n_scatters"]
        total_jumps += o["n_scatters"]
        m = o["n_scatters"]
        assert np.array_equal(g["jump_ch"][i, :m], [(ie * 12 + ig) * 3 + q for (_, _, ie, ig, q) in o["jumps"]])
        assert np.allclose(g["jump_t"][i, :m], [t for (_, t, _, _, _) in o["jumps"]], rtol=0, atol=1e-12)
        assert g["n_saved"][i] == o["n_saved"] == 16
        assert np.max(np.abs(g["pop"][i] - o["pop"])) < 1e-9
        assert np.max(np.abs(g["force"][i] - o["F"])) < 1e-9 * max(1.0, np.max(np.abs(o["F"])))
        assert np.max(np.abs(g["r"][i] - o["r"])) < 1e-9 * max(1.0, np.max(np.abs(o["r"])))
        assert np.max(np.abs(g["v"][i] - o["v"])) < 1e-12 + 1e-9 * np.max(np.abs(o["v"]))
        assert np.max(np.abs(g["psi_final"][i] - o["psi"])) < 1e-9
        assert np.max(np.abs(g["force_sum"][i] - o["force_sum"])) < 1e-9 * max(1.0, np.max(np.abs(o["force_sum"])))
        assert g["n_avg"][i] == o["n_avg"] == 1500
    assert total_jumps >= 3


def test_escape_stops_trajectory(mol, dev):
    prob, tab = _problem(mol, kick=False)
    r0 = np.array([[0.0, 0.0, 1000.0], [0.0, 0.0, 0.0]])
    v0 = np.array([[0.0, 0.0, 500.0], [0.0, 0.0, 0.0]])
    _, g = _run_gpu(prob, r0, v0, None, 300, 10, 5, "fp64", r_escape=2000.0)
    o = mc.run(tab, r0[0], v0[0], np.eye(16)[0], 0.01, 300, 10, seed=5, traj=0, shift=mc.default_shift(tab),
               r_escape=2000.0)
    assert g["escaped"].tolist() == [1, 0] and o["escaped"] == 1
    assert g["n_saved"][0] == o["n_saved"] < 31 and g["n_saved"][1] == 31
    assert np.max(np.abs(g["r"][0, :o["n_saved"]] - o["r"][:o["n_saved"]])) < 1e-9 * 3000


def test_fp32_before_first_jump(mol, dev):
    prob, tab = _problem(mol, s0=1.0)
    r0 = np.array([[300.0, -200.0, 100.0], [-50.0, 800.0, 0.0]])
    v0 = np.array([[0.05, 0.0, 0.1], [0.0, -0.1, 0.0]])
    _, g = _run_gpu(prob, r0, v0, None, 300, 20, 99, "fp32", jump_cap=8)
    for i in range(2):
        o = mc.run(tab, r0[i], v0[i], np.eye(16)[0], 0.01, 300, 20, seed=99, traj=i, shift=mc.default_shift(tab))
        first = o["jumps"][0][0] if o["jumps"] else 300
        ns = first // 20 + 1
        assert ns >= 5
        assert np.max(np.abs(g["pop"][i, :ns] - o["pop"][:ns])) < 2e-5
        fs = max(1e-3, np.max(np.abs(o["F"][:ns])))
        assert np.max(np.abs(g["force"][i, :ns] - o["F"][:ns])) < 2e-5 * fs + 2e-6


@pytest.mark.parametrize("v_ms", [0.0, 0.5, 1.0])
def test_ensemble_force_equals_master_equation(mol, dev, v_ms):
    """Fixed velocity along z (mass x 1e12, no recoil), exact waiting-time
    jumps: E[<F>/<psi|psi>](t) from 16,384 FP32 trajectories vs Tr(rho(t) F(t))
    of the master equation this MCWF unravels, at 31 save points over
    t = 30 / Gamma, within 5 standard errors + 0.5 % of the force scale.
    The reference jump rule (1 - |psi|^2 > -ln U) is also run and must stay
    within 5 % (it under-samples jumps, see molmot_amd/obe/stochastic.py)."""
    prob, tab = _problem(mol, s0=2.0, kick=False, heavy=True)
    n = 16384
    vn = v_ms / prob.v_nat
    r0 = np.tile([500.0, 300.0, 200.0], (n, 1))
    v0 = np.tile([0.0, 0.0, vn], (n, 1))
    rho0 = np.zeros((16, 16), dtype=complex)
    rho0[0, 0] = 1.0
    for rule, rel in (("exact", 0.005), ("reference", 0.05)):
        plan, g = _run_gpu(prob, r0, v0, None, 3000, 100, 2024 + int(10 * v_ms), "fp32", jump_rule=rule)
        if rule == "exact":
            _, F, _ = mc.lindblad_path(tab, r0[0], v0[0], rho0, 0.01, 3000, plan.shift)
            Fl = F[::100]
        Fm = g["force"].mean(0)
        se = g["force"].std(0) / np.sqrt(n)
        scale = np.max(np.abs(Fl))
        bad = np.abs(Fm - Fl) > 5 * se + rel * scale
        print(rule, "max dev / scale", np.max(np.abs(Fm - Fl)) / scale, "max se / scale", np.max(se) / scale,
              flush=True)
        assert not bad.any(), np.argwhere(bad)
        assert g["n_scatters"].mean() > 0.05


def test_drop_in_solver_api(mol, dev):
    """SSESolver.run / run_ensemble shapes (tests/test_all_functions.py:432-450
    of the reference: populations of a short trajectory sum to 1)."""
    from molmot_amd.obe.stochastic import SSEProblem, SSESolver, run_ensemble
    bp = make_beam_pairs_dc4(mol, -0.30, 0.40, 1.0)
    prob = SSEProblem(mol, bp, B_gradient=0.16)
    res = SSESolver(prob).run(np.zeros(3), np.zeros(3), t_max=1e-7, save_every=10, rng_seed=42)
    assert res.times.shape[0] > 0 and res.positions.shape[1] == 3
    assert np.all(np.abs(res.populations.sum(1) - 1.0) < 0.05)
    assert res.forces.shape == res.positions.shape
    out = run_ensemble(prob, 5, lambda: np.zeros(3), lambda: np.array([0, 0, 0.5]), 2e-7, save_every=20,
                       rng_seed=3)
    assert len(out) == 5 and all(o.velocities.shape == out[0].velocities.shape for o in out)
    assert abs(out[0].velocities[0, 2] - 0.5) < 1e-12
