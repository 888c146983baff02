"""GPU parity of the rate-equation force kernels against the FP64 oracle and
the reference fixtures.  Tolerances: the north-star bar is 1e-10 relative in
FP64 (norm-relative, max|a-b| / max|b|, since F crosses zero); the kernel
keeps the reference's detuning arithmetic exactly, so the observed error is
~1e-15 and the tests assert 1e-12 (force, scattering rate) and 1e-13
absolute on populations."""
import numpy as np
import pytest

import oracle
from molmot_amd.constants import HBAR_LEGACY, hbar as HBAR
from molmot_amd.molecules import synthetic_molecule
from molmot_amd.obe.fields import LaserBeam, beam_table, make_dc_mot_beams, make_hallas_beams, hallas_saturations

pytestmark = pytest.mark.gpu

B_HALLAS = 21.4 * 1e-4 / 1e-2


def maxrel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def gpu_points(mol, bt, v, z, B, hbar, Geff=1.0):
    import torch
    from molmot_amd.hip import HIPMolData, rate_eq_points_device
    plan = HIPMolData(mol, bt, Geff, hbar=hbar)
    dev = torch.device("cuda", 0)
    F, P, R = rate_eq_points_device(plan, torch.as_tensor(np.asarray(v, dtype=np.float64)).to(dev),
                                    torch.as_tensor(np.asarray(z, dtype=np.float64)).to(dev), B)
    torch.cuda.synchronize()
    return F.cpu().numpy(), P.cpu().numpy(), R.cpu().numpy()


def test_three_way_fixture(mol, golden, torch_dev):
    """GPU vs the reference's own CUDA column (legacy hbar) and NumPy column
    (CODATA hbar) of validation/three_way_results.npz."""
    from molmot_amd.hip import force_vs_z_hip, force_vs_v_hip, spring_constant_hip, damping_coefficient_hip
    tw = golden("three_way_results.npz")
    beams, _, _ = make_hallas_beams(mol)
    unit = HBAR * mol.k * mol.Gamma / 2.0
    for hb, key in [(HBAR_LEGACY, "gpu"), (HBAR, "cpu_s")]:
        Fz, Pz, Rz = force_vs_z_hip(mol, beams, tw["z_arr"], 0.0, B_HALLAS, hbar=hb)
        Fv = force_vs_v_hip(mol, beams, tw["v_arr"], 0.0, B_HALLAS, hbar=hb)
        assert maxrel(Fz / unit, tw["Fz_" + key]) < 1e-12
        assert maxrel(Fv / unit, tw["Fv_" + key]) < 1e-12
        assert np.max(np.abs(Pz - tw["Pz_" + key])) < 1e-13
        k = spring_constant_hip(mol, beams, B_HALLAS, hbar=hb)
        beta = damping_coefficient_hip(mol, beams, B_HALLAS, hbar=hb) / mol.mass
        assert abs(k / tw["k_" + key] - 1) < 1e-10
        assert abs(beta / tw["beta_" + key] - 1) < 1e-10


@pytest.mark.parametrize("cfg", ["hallas5mm", "dc", "dc_strong"])
def test_map_vs_oracle(mol, torch_dev, cfg):
    """256 x 256 (v, z) map vs the oracle's JIT flavour, same hbar."""
    if cfg == "hallas5mm":
        beams = make_hallas_beams(mol, sats=hallas_saturations(mol.Gamma, beam_radius_m=5e-3))[0]
        B = B_HALLAS
    elif cfg == "dc":
        beams, B = make_dc_mot_beams(mol, -0.2, 0.7, 1.0), 0.16
    else:
        beams, B = make_dc_mot_beams(mol, -2.0, 0.3, 100.0), 0.5
    bt = beam_table(beams)
    from molmot_amd.hip import force_map_2d_hip
    v = np.linspace(-30, 30, 256)
    z = np.linspace(-10e-3, 10e-3, 256)
    F, P, R = force_map_2d_hip(mol, beams, v, z, B, hbar=HBAR)
    Fo, Po, Ro = oracle.rate_eq_map(mol, bt, v, z, B, hbar=HBAR, flavour="jit")
    assert maxrel(F, Fo) < 1e-12
    assert maxrel(R, Ro) < 1e-12
    assert np.max(np.abs(P - Po)) < 1e-13
    assert np.max(np.abs(P.sum(-1) - 1.0)) < 1e-14
    assert np.all(P >= 0)


def test_full_size_map_properties(mol, torch_dev):
    """1024 x 1024 map (z in [-10,10] mm, v in [-30,30] m/s): populations sum
    to 1 everywhere, 4096 random grid points match the oracle, and the
    symmetric DC configuration gives F(-v,-z) = -F(v,z)."""
    import torch
    from molmot_amd.hip import HIPMolData, rate_eq_map_device
    beams = make_hallas_beams(mol, sats=hallas_saturations(mol.Gamma, beam_radius_m=5e-3))[0]
    bt = beam_table(beams)
    plan = HIPMolData(mol, bt, 1.0, hbar=HBAR)
    v = np.linspace(-30, 30, 1024)
    z = np.linspace(-10e-3, 10e-3, 1024)
    F, P, R = rate_eq_map_device(plan, torch.from_numpy(v).to(torch_dev), torch.from_numpy(z).to(torch_dev),
                                 B_HALLAS)
    psum_err = float((P.sum(1) - 1.0).abs().max())
    assert psum_err < 1e-14
    assert bool((P >= 0).all())
    F = F.cpu().numpy().reshape(1024, 1024)
    R = R.cpu().numpy().reshape(1024, 1024)
    rng = np.random.default_rng(1)
    iv = rng.integers(0, 1024, 4096)
    iz = rng.integers(0, 1024, 4096)
    Fo, _, Ro = oracle.rate_eq_points(mol, bt, v[iv], z[iz], B_HALLAS, hbar=HBAR, flavour="jit")
    assert float(np.max(np.abs(F[iv, iz] - Fo))) / float(np.max(np.abs(F))) < 1e-12
    assert maxrel(R[iv, iz], Ro) < 1e-12


def test_antisymmetry_dc(mol, torch_dev):
    """The DC configuration is symmetric under (v, z) -> (-v, -z) with the
    beam pair swapped, so F is odd; checks the whole grid at rounding level."""
    from molmot_amd.hip import force_map_2d_hip
    beams = make_dc_mot_beams(mol, -0.3, 0.4, 1.0)
    v = np.linspace(-5, 5, 101)
    z = np.linspace(-4e-3, 4e-3, 101)
    F, _, _ = force_map_2d_hip(mol, beams, v, z, 0.16, hbar=HBAR)
    Fo, _, _ = oracle.rate_eq_map(mol, beam_table(beams), v, z, 0.16, hbar=HBAR, flavour="jit")
    # the reference itself is only approximately odd (the split is not
    # symmetric in the molecule's structure); compare the asymmetry with the
    # oracle's instead of assuming it vanishes
    asym = F + F[::-1, ::-1]
    asym_o = Fo + Fo[::-1, ::-1]
    assert np.max(np.abs(asym - asym_o)) < 1e-12 * np.max(np.abs(F))


@pytest.mark.parametrize("ng,ne,seed", [(36, 4, 3), (4, 2, 5), (3, 1, 2), (2, 1, 7)])
def test_synthetic_sizes(torch_dev, ng, ne, seed):
    m = synthetic_molecule(ng, ne, seed=seed)
    bt = beam_table(make_dc_mot_beams(m, -0.5, 0.4, 2.0))
    rng = np.random.default_rng(seed)
    v = rng.uniform(-10, 10, 2000)
    z = rng.uniform(-5e-3, 5e-3, 2000)
    F, P, R = gpu_points(m, bt, v, z, 0.2, HBAR)
    Fo, Po, Ro = oracle.rate_eq_points(m, bt, v, z, 0.2, hbar=HBAR, flavour="jit")
    assert maxrel(F, Fo) < 1e-12
    assert maxrel(R, Ro) < 1e-12
    assert np.max(np.abs(P - Po)) < 1e-12


def test_dark_state_fallback(mol, torch_dev):
    """Corner cases the reduced solve does not cover run the restated
    pivot-skip elimination: all s0 = 0 (every ground state dark) gives
    p = e0 and F = R = 0 as the reference JIT path does; a single sigma+
    beam leaves some ground states uncoupled."""
    bt = beam_table(make_dc_mot_beams(mol, -0.2, 0.7, 0.0))
    F, P, R = gpu_points(mol, bt, [0.0, 1.0], [1e-3, -2e-3], 0.16, HBAR)
    assert np.all(F == 0.0) and np.all(R == 0.0)
    assert np.all(P[:, 0] == 1.0) and np.all(P[:, 1:] == 0.0)
    one = beam_table([LaserBeam(1.0, mol.omega_J32 - 3e6, 2, 3.0)])
    v = np.linspace(-3, 3, 50)
    z = np.linspace(-3e-3, 3e-3, 50)
    F, P, R = gpu_points(mol, one, v, z, 0.16, HBAR)
    Fo, Po, Ro = oracle.rate_eq_points(mol, one, v, z, 0.16, hbar=HBAR, flavour="jit")
    assert np.max(np.abs(F - Fo)) <= 1e-12 * max(np.max(np.abs(Fo)), 1e-300)
    assert np.max(np.abs(P - Po)) < 1e-12


def test_empty_and_single(mol, torch_dev):
    import torch
    from molmot_amd.hip import HIPMolData, rate_eq_map_device
    plan = HIPMolData(mol, make_dc_mot_beams(mol), 1.0)
    e = torch.empty(0, dtype=torch.float64, device=torch_dev)
    one = torch.zeros(1, dtype=torch.float64, device=torch_dev)
    F, P, R = rate_eq_map_device(plan, e, one, 0.16)
    assert F.numel() == 0
    F, P, R = rate_eq_map_device(plan, one, one, 0.16)
    assert F.numel() == 1 and abs(float(P.sum()) - 1.0) < 1e-14


def test_simulator_api(mol, torch_dev):
    """Drop-in API: DCMOTSimulator / force_scan functions vs the oracle."""
    from molmot_amd.mot import DCMOTSimulator, RFMOTSimulator, force_vs_z, force_vs_v, spring_constant, \
        damping_coefficient
    sim = DCMOTSimulator(mol, delta_Gamma=-0.30, split_Gamma=0.40, s0=1.0, B_gradient=0.16)
    bt = beam_table(sim._beams)
    F, p, R = sim.force(0.5, 1e-3)
    Fo, po, Ro = oracle.rate_eq(mol, bt, 0.5, 1e-3, 0.16, hbar=HBAR, flavour="numpy")
    assert abs(F - Fo) <= 1e-12 * abs(Fo) and np.max(np.abs(p - po)) < 1e-13
    z = np.linspace(-3e-3, 3e-3, 5)
    Fz, Pz, Rz = force_vs_z(sim, z)
    assert Fz.shape == (5,) and Pz.shape == (5, 12)
    assert force_vs_v(sim, np.linspace(-5, 5, 5)).shape == (5,)
    assert spring_constant(sim) > 0 and damping_coefficient(sim) > 0
    rf = RFMOTSimulator(mol, delta_Gamma=-1.0, s0=1.0, B_gradient=0.16)
    assert rf.force(0.0, 1e-3)[0] < 0
