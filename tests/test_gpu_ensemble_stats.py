"""On-device ensemble statistics (mm_traj_stats_1d) against the host
functions of molmot/mot/analysis.py:128-288 applied to stored trajectories:
the FP64 oracle's (no kicks) and the GPU integrator's own saved ones (with
Philox kicks: the segmented run must reproduce the unbroken one)."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(mol):
    from molmot_amd.mot import DCMOTSimulator
    sim = DCMOTSimulator(mol, -0.2, 0.7, 1.0, 0.16)
    rng = np.random.default_rng(23)
    n = 700
    z0 = rng.normal(0, 2e-3, n)
    v0 = rng.normal(0, 3.0, n) + rng.uniform(-8, 8, n)
    return sim, z0, v0


def _host(res, mass, nb):
    from molmot_amd.mot import analysis as an
    tc, sz, sv, na = an.sigma_vs_time(res, nb)
    _, T = an.temperature_vs_time(res, mass, nb)
    _, n0 = an.density_vs_time(res, 2000, nb)
    return tc, sz, sv, na, T, n0


def _device(sim, z0, v0, t_max, dt, nb, **kw):
    from molmot_amd.mot import analysis as an
    tc, sz, sv, na = an.sigma_vs_time_ensemble(sim, z0, v0, t_max, dt, nb, **kw)
    _, T = an.temperature_vs_time_ensemble(sim, z0, v0, t_max, dt, nb, **kw)
    _, n0 = an.density_vs_time_ensemble(sim, z0, v0, t_max, dt, 2000, nb, **kw)
    return tc, sz, sv, na, T, n0


def test_stats_vs_oracle_trajectories(setup, mol):
    from molmot_amd.constants import hbar
    from molmot_amd.obe.fields import beam_table
    sim, z0, v0 = setup
    dt, n_steps, nb = 1e-5, 2500, 37
    print(f"ensemble stats: {z0.size} particles x {n_steps} samples vs oracle", flush=True)
    _, _, na, zt, vt = oracle.traj_euler(mol, beam_table(sim._beams), z0, v0, dt, n_steps, sim.B_gradient,
                                         hbar=sim._hip_plan.hbar_val, save_full=True)
    t = np.arange(n_steps) * dt
    res = [(t, zt[i], vt[i]) for i in range(z0.size)]
    h = _host(res, mol.mass, nb)
    g = _device(sim, z0, v0, n_steps * dt + 0.5 * dt, dt, nb)
    assert np.array_equal(h[0], g[0]) and np.array_equal(h[3], g[3])
    assert 0 < h[3].min() < z0.size          # some molecules leave
    for a, b in zip(h[1:], g[1:]):
        assert np.allclose(a, b, rtol=1e-9, atol=0, equal_nan=True)


def test_stats_with_kicks_vs_saved_gpu_trajectories(setup, mol):
    from molmot_amd.hip import traj_euler_device, recoil_kick_scale
    sim, z0, v0 = setup
    dt, n_steps, nb = 1e-5, 1801, 900    # bins of two steps: segments ending at odd and even steps
    ks = 40 * recoil_kick_scale(mol)
    print(f"ensemble stats: {z0.size} particles x {n_steps} samples, kicks", flush=True)
    dev = torch.device("cuda", 0)
    zf, vf, na, zt, vt = traj_euler_device(sim._hip_plan, torch.from_numpy(z0).to(dev), torch.from_numpy(v0).to(dev),
                                           sim.B_gradient, dt, n_steps, save_full=True, kick_scale=ks, seed=99)
    zt, vt = zt.cpu().numpy(), vt.cpu().numpy()
    t = np.arange(n_steps) * dt
    res = [(t, zt[i], vt[i]) for i in range(z0.size)]
    h = _host(res, mol.mass, nb)
    g = _device(sim, z0, v0, n_steps * dt + 0.5 * dt, dt, nb, kick_scale=ks, seed=99)
    assert np.array_equal(h[3], g[3])
    for a, b in zip(h[1:], g[1:]):
        assert np.allclose(a, b, rtol=1e-12, atol=0, equal_nan=True)
