"""The work queue's give-up path, driven from the host: with a wait bound of
0 s a wavefront that has to wait for a group's next chunk gives up at once,
every wavefront exits normally, the status kernel sets the counts to -1 and
the Python wrappers raise.  The ensemble is one group of 64 particles over
16 chunks on a persistent grid of many more wavefronts, so the wavefront that
claims chunk 1 always finds it unpublished.  The default bound is restored
and the same launch then matches one thread per particle bit for bit."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(mol):
    from molmot_amd.hip import HIPMolData
    from molmot_amd.obe.fields import make_dc_mot_beams
    from molmot_amd.constants import hbar
    plan = HIPMolData(mol, make_dc_mot_beams(mol, -0.2, 0.7, 1.0), 1.0, hbar=hbar)
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(3)
    z0 = torch.from_numpy(rng.normal(0, 1e-3, 64)).to(dev)
    v0 = torch.from_numpy(rng.uniform(-3, 3, 64)).to(dev)
    return plan, z0, v0


def test_queue_gives_up_and_raises(setup):
    from molmot_amd.hip import _lib
    from molmot_amd.hip.trajectories import traj_euler_device
    plan, z0, v0 = setup
    n_steps = 2049          # 2048 updates: 16 chunks of 128
    print("queue give-up: 64 particles x 2049 samples, wait bound 0 s", flush=True)
    _lib.set_queue_wait(0.0)
    try:
        zf, vf, na, _, _ = traj_euler_device(plan, z0, v0, 0.16, 1e-5, n_steps, schedule="queue", check=False)
        torch.cuda.synchronize()
        assert (na.cpu().numpy() == -1).all()
        with pytest.raises(_lib.MolmotHIPError, match="timed out"):
            traj_euler_device(plan, z0, v0, 0.16, 1e-5, n_steps, schedule="queue")
    finally:
        _lib.set_queue_wait(None)
    torch.cuda.synchronize()
    zq, vq, nq, _, _ = traj_euler_device(plan, z0, v0, 0.16, 1e-5, n_steps, schedule="queue")
    zg, vg, ng, _, _ = traj_euler_device(plan, z0, v0, 0.16, 1e-5, n_steps, schedule="grid")
    assert torch.equal(zq, zg) and torch.equal(vq, vg) and torch.equal(nq, ng)
    assert (ng.cpu().numpy() > 0).all()


def test_tangent_and_temperature_queue_give_up(setup, mol):
    from molmot_amd.hip import _lib
    from molmot_amd.adjoint import ForceProblem, trajectory_tangent
    from molmot_amd.mot import DCMOTSimulator
    from molmot_amd.propagation import monte_carlo_temperature
    from molmot_amd.obe.fields import make_dc_mot_beams
    plan, z0, v0 = setup
    prob = ForceProblem(mol, make_dc_mot_beams(mol, -0.2, 0.7, 1.0), hbar=plan.hbar_val)
    theta = prob.theta0(0.16)
    print("queue give-up: tangent 64 x 2048 updates and temperature 100 x 4000 steps, bound 0 s", flush=True)
    _lib.set_queue_wait(0.0)
    try:
        with pytest.raises(_lib.MolmotHIPError, match="timed out"):
            trajectory_tangent(prob, theta, z0, v0, mol.mass, 1e-5, 2048, schedule="queue")
        sim = DCMOTSimulator(mol, -0.2, 0.7, 1.0, 0.16)
        with pytest.raises(_lib.MolmotHIPError, match="timed out"):
            monte_carlo_temperature(sim, n_particles=100, t_max=4e-3, dt=1e-6, t_equilibrium=1e-3,
                                    schedule="queue")
    finally:
        _lib.set_queue_wait(None)
    torch.cuda.synchronize()
    o = trajectory_tangent(prob, theta, z0, v0, mol.mass, 1e-5, 2048, schedule="queue")
    assert (o["n_done"].cpu().numpy() > 0).all()
