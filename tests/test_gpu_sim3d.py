"""GPU 3D rate-equation force (in-LDS Jacobi) and leapfrog vs the FP64 oracle.

* Non-degenerate Zeeman spectra (seeded synthetic 12+4 molecule, relative
  gaps >= 1e-2): F, R and populations equal the oracle's restatement of
  rate_eq_force_3d (numpy eigh, LU) to 1e-10 relative.
* r = 0 (B = 0, H_Z = 0, identity eigenbasis in both): SrOH at several
  velocities, same bar.
* SrOH elsewhere (six-fold degenerate eigenvalues, the reference's force is
  basis dependent): the reference's own checks (F(0) ~ 0, restoring F_z at
  z = 1 mm) and basis-independent invariants (p >= 0, sum p = 1, F and R
  finite, R >= 0).
* Two-level saturated closed form.
* Leapfrog: trajectories (stored r, v every step, escape fill) equal the
  oracle's simulate_trajectory_3d restatement to 1e-10 relative.
* Spin-rotation basis (the default for spin-1/2 Zeeman tables: SrOH, built
  CaF): F, R, p equal the oracle's restatement of the same basis convention
  (oracle.sim3d basis="spin") to 1e-9 relative at random points, on the
  +-z axes and at B = 0; the leapfrog likewise.  The kernel takes branching
  ratios over the unitary-invariant column sums and solves the excited chain
  by state reduction, the oracle follows the reference (per-point sums, LU),
  so they differ by the tables' own 1e-11 departure from the spin algebra
  and rounding.  The Jacobi path remains selectable (basis="jacobi") and
  gives the same basis-independent invariants."""
import numpy as np
import pytest

from oracle import sim3d as s3
from molmot_amd.molecules import synthetic_molecule
from molmot_amd.molecules.sroh import MolecularData
from molmot_amd.obe.beam_pairs import make_beam_pairs_dc4

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def _sim(m, pairs, B=0.16, basis="auto"):
    from molmot_amd.mot.simulator_3d import MOTSimulator3D
    return MOTSimulator3D(m, pairs, B, basis=basis)


def _obeams(sim):
    return [dict(k_hat=b.k_hat, freq=b.freq, q=b.q_lab, s0=b.s0, w=b.beam_radius) for b in sim.beams]


def test_force_synthetic_vs_oracle(dev):
    m = synthetic_molecule(12, 4, seed=3)
    sim = _sim(m, make_beam_pairs_dc4(m, -0.8, 0.5, 2.0))
    rng = np.random.default_rng(1)
    r = rng.uniform(-4e-3, 4e-3, size=(24, 3))
    v = rng.uniform(-3.0, 3.0, size=(24, 3))
    F, R, p = sim.force_3d_batch(r, v)
    ob = _obeams(sim)
    Fo = np.zeros_like(F)
    for i in range(len(r)):
        Fo[i], Ro, po, _ = s3.force(m, ob, r[i], v[i], 0.16)
        assert abs(R[i] - Ro) < 1e-10 * Ro
        assert np.max(np.abs(p[i] - po)) < 1e-10
    scale = np.max(np.abs(Fo))
    assert np.max(np.abs(F - Fo)) < 1e-10 * scale


def test_force_sroh_at_centre_vs_oracle(mol, dev):
    sim = _sim(mol, make_beam_pairs_dc4(mol, -0.30, 0.40, 1.0))
    v = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 1.5], [1.0, -2.0, 0.5], [-3.0, 0.2, 0.0]])
    F, R, p = sim.force_3d_batch(np.zeros((4, 3)), v)
    ob = _obeams(sim)
    for i in range(4):
        Fo, Ro, po, _ = s3.force(mol, ob, np.zeros(3), v[i], 0.16)
        assert np.max(np.abs(F[i] - Fo)) < 1e-10 * max(np.max(np.abs(Fo)), 1e-22)
        assert abs(R[i] - Ro) < 1e-10 * Ro
        assert np.max(np.abs(p[i] - po)) < 1e-10


def test_sroh_reference_checks_and_invariants(mol, dev):
    """tests/test_all_functions.py:534-555 plus invariants over a grid."""
    sim = _sim(mol, make_beam_pairs_dc4(mol, -0.30, 0.40, 1.0))
    F0, R0, p0 = sim.force_3d(np.zeros(3), np.zeros(3))
    assert F0.shape == (3,) and np.linalg.norm(F0) < 1e-20
    F1, _, _ = sim.force_3d(np.array([0.0, 0.0, 1e-3]), np.zeros(3))
    assert F1[2] < 0
    rng = np.random.default_rng(2)
    r = rng.uniform(-5e-3, 5e-3, size=(256, 3))
    v = rng.uniform(-5, 5, size=(256, 3))
    F, R, p = sim.force_3d_batch(r, v)
    assert np.all(np.isfinite(F)) and np.all(R >= 0)
    assert np.max(np.abs(p.sum(1) - 1)) < 1e-13 and p.min() >= 0


def test_two_level_closed_form(dev):
    from molmot_amd.mot.simulator_3d import Beam3D, rate_eq_force_3d_batch
    tdm = np.zeros((2, 2, 3), dtype=complex)
    tdm[0, 1, 2] = tdm[1, 0, 2] = 1.0
    Z = np.zeros((2, 2), dtype=complex)
    m = MolecularData(energies=np.array([1.0e9, 4.0e14]), tdm=tdm, zeeman_x=Z, zeeman_y=Z, zeeman_z=Z,
                      n_ground=1, n_excited=1, n_states=2)
    G = m.Gamma
    f = m.energies[1] - m.energies[0] - 0.7 * G / (2 * np.pi)
    beams = [Beam3D([0, 0, 1], f, 2, 1.6)]
    vz = np.array([0.0, 1.0, -2.0])
    F, R, _ = rate_eq_force_3d_batch(m, beams, np.tile([0, 0, 1e-3], (3, 1)), np.c_[np.zeros((3, 2)), vz], 0.2)
    delta = 2 * np.pi * (f - m.k * vz / (2 * np.pi) - (m.energies[1] - m.energies[0]))
    L = (G / 2) ** 2 / (delta ** 2 + (G / 2) ** 2)
    rate = G / 2 * 1.6 * L / (1 + 1.6 * L)
    assert np.allclose(R, rate, rtol=1e-12, atol=0)
    assert np.allclose(F[:, 2], s3.HBAR * m.k * rate, rtol=1e-12, atol=0)


def test_leapfrog_vs_oracle(dev):
    m = synthetic_molecule(12, 4, seed=3)
    sim = _sim(m, make_beam_pairs_dc4(m, -0.8, 0.5, 2.0))
    r0 = np.array([[1e-3, -2e-3, 5e-4], [0.0, 0.0, 3e-3], [2e-3, 2e-3, -1e-3], [0.0, 0.0, 4.5e-3]])
    v0 = np.array([[0.5, 0.0, -1.0], [0.0, 0.0, 2.0], [-1.0, 0.5, 0.0], [0.0, 0.0, 30.0]])
    n_steps, dt, rsc = 150, 1e-5, 0.006
    t, r, v = sim._ensemble(r0, v0, n_steps * dt, dt, rsc)
    ob = _obeams(sim)
    for i in range(4):
        print(f"particle {i}", flush=True)
        to, ro, vo = s3.leapfrog(m, ob, r0[i], v0[i], n_steps * dt, dt, 0.16, r_escape=rsc)
        assert np.max(np.abs(r[i] - ro)) < 1e-10 * np.max(np.abs(ro))
        assert np.max(np.abs(v[i] - vo)) < 1e-10 * np.max(np.abs(vo))
    assert np.linalg.norm(r[3, -1]) > rsc


def test_force_map_shape(mol, dev):
    from molmot_amd.mot.simulator_3d import compute_3d_force_map
    sim = _sim(mol, make_beam_pairs_dc4(mol, -0.30, 0.40, 1.0))
    c1, c2, Fa, Fb, R = compute_3d_force_map(sim, "xz", n_points=9, extent=4e-3)
    assert Fa.shape == Fb.shape == R.shape == (9, 9)
    F, Rr, _ = sim.force_3d(np.array([c1[2], 0.0, c2[7]]), np.zeros(3))
    assert abs(Fa[2, 7] - F[0]) == 0 and abs(Fb[2, 7] - F[2]) == 0 and R[2, 7] == Rr


def _caf():
    from molmot_amd.molecules import builders as bl
    return bl.zeeman_per_gauss(bl.load_caf_from_sroh_structure())


@pytest.mark.parametrize("which", ["sroh", "caf"])
def test_spin_basis_vs_oracle(mol, dev, which):
    m = mol if which == "sroh" else _caf()
    sim = _sim(m, make_beam_pairs_dc4(m, -0.9, 0.6, 1.7))
    assert sim.plan.basis == "spin"
    rng = np.random.default_rng(8)
    r = rng.uniform(-5e-3, 5e-3, size=(40, 3))
    v = rng.uniform(-4.0, 4.0, size=(40, 3))
    r[0] = 0.0
    r[1] = [0.0, 0.0, 2e-3]
    r[2] = [0.0, 0.0, -2e-3]
    r[3] = [1e-12, 0.0, -2e-3]
    F, R, p = sim.force_3d_batch(r, v)
    ob = _obeams(sim)
    Fo = np.zeros_like(F)
    for i in range(len(r)):
        Fo[i], Ro, po, _ = s3.force(m, ob, r[i], v[i], 0.16, basis="spin")
        assert abs(R[i] - Ro) < 1e-9 * Ro, i
        assert np.max(np.abs(p[i] - po)) < 1e-9, i
    assert np.max(np.abs(F - Fo)) < 1e-9 * np.max(np.abs(Fo))
    assert np.max(np.abs(p.sum(1) - 1)) < 1e-12 and p.min() >= 0


def test_spin_and_jacobi_paths_share_invariants(mol, dev):
    """SrOH on both paths: the reference's own checks (F(0) ~ 0, restoring
    F_z at z = 1 mm), equal force at B = 0 (both use the identity basis
    there), and p >= 0, sum p = 1, R >= 0 everywhere."""
    pairs = make_beam_pairs_dc4(mol, -0.30, 0.40, 1.0)
    sj, ss = _sim(mol, pairs, basis="jacobi"), _sim(mol, pairs, basis="spin")
    assert sj.plan.basis == "jacobi" and ss.plan.basis == "spin"
    rng = np.random.default_rng(12)
    r = rng.uniform(-5e-3, 5e-3, size=(64, 3))
    v = rng.uniform(-5, 5, size=(64, 3))
    r[0] = 0.0
    for sim in (sj, ss):
        F0, _, _ = sim.force_3d(np.zeros(3), np.zeros(3))
        assert np.linalg.norm(F0) < 1e-20
        F1, _, _ = sim.force_3d(np.array([0.0, 0.0, 1e-3]), np.zeros(3))
        assert F1[2] < 0
        F, R, p = sim.force_3d_batch(r, v)
        assert np.all(np.isfinite(F)) and np.all(R >= 0)
        assert np.max(np.abs(p.sum(1) - 1)) < 1e-12 and p.min() >= 0
    Fj = sj.force_3d_batch(r[:1], v[:1])[0]
    Fs = ss.force_3d_batch(r[:1], v[:1])[0]
    assert np.max(np.abs(Fj - Fs)) < 1e-10 * np.max(np.abs(Fj))


def test_spin_leapfrog_vs_oracle(mol, dev):
    sim = _sim(mol, make_beam_pairs_dc4(mol, -0.9, 0.6, 1.7))
    r0 = np.array([[1e-3, -2e-3, 5e-4], [0.0, 0.0, 3e-3], [2e-3, 2e-3, -1e-3], [0.0, 0.0, 4.5e-3]])
    v0 = np.array([[0.5, 0.0, -1.0], [0.0, 0.0, 2.0], [-1.0, 0.5, 0.0], [0.0, 0.0, 30.0]])
    n_steps, dt, rsc = 120, 1e-5, 0.006
    t, r, v = sim._ensemble(r0, v0, n_steps * dt, dt, rsc)
    ob = _obeams(sim)
    for i in range(4):
        print(f"spin leapfrog particle {i}", flush=True)
        to, ro, vo = s3.leapfrog(mol, ob, r0[i], v0[i], n_steps * dt, dt, 0.16, r_escape=rsc, basis="spin")
        assert np.max(np.abs(r[i] - ro)) < 1e-9 * np.max(np.abs(ro))
        assert np.max(np.abs(v[i] - vo)) < 1e-9 * np.max(np.abs(vo))
    assert np.linalg.norm(r[3, -1]) > rsc


def test_segmented_leapfrog_matches_one_launch(mol, dev):
    """The segmented run (compaction of the particles still inside, 1, 4 or
    16 lanes per particle) against one launch of the same steps: equal up
    to the order of the per-ground sums (1e-9), identical escape steps; as
    a single segment of one lane per particle, bit for bit."""
    import torch
    from molmot_amd import configs
    from molmot_amd.mot.simulator_3d import Force3DPlan, leapfrog3d_device, make_6beam_config
    beams = make_6beam_config(mol, make_beam_pairs_dc4(mol, -1.0, 0.5, 2.0), 5e-3)
    plan = Force3DPlan(mol, beams)
    assert plan.basis == "spin"
    # 40,000 particles: more than the 32,768 that 4 lanes each would take on
    # 1,024 SIMDs, so the first segment runs one lane per particle
    n, ns = 40000, 600
    r0h, v0h = configs.ensemble_initial_conditions_3d(mol, n)
    r0, v0 = torch.from_numpy(r0h).to(dev), torch.from_numpy(v0h).to(dev)
    print("one launch", flush=True)
    a = leapfrog3d_device(plan, r0, v0, 0.214, 1e-5, ns, save_full=True, segmented=False)
    print("segmented, 64-step segments", flush=True)
    b = leapfrog3d_device(plan, r0, v0, 0.214, 1e-5, ns, save_full=True, segmented=True, seg_steps=64)
    print("segmented, one segment", flush=True)
    c = leapfrog3d_device(plan, r0, v0, 0.214, 1e-5, ns, save_full=True, segmented=True, seg_steps=ns + 5)
    torch.cuda.synchronize()
    na, nb_, nc = (x["n_actual"] for x in (a, b, c))
    alive = int((na == ns).sum())
    print("alive at the end", alive, flush=True)
    assert 0 < alive < n
    assert torch.equal(na, nb_) and torch.equal(na, nc)
    for k in ("r_traj", "v_traj", "r_final", "v_half_final"):
        assert torch.equal(a[k], c[k]), k
        dev_ = float((a[k] - b[k]).abs().max())
        assert dev_ <= 1e-9 * float(a[k].abs().max()), (k, dev_)
    plan.close()
