"""GPU Lindblad RK4 propagators vs the FP64 numpy oracle.

Bars: FP64 path 1e-10 (force, norm-relative over the batch) and 1e-12 on
rho; FP32 paths (VALU and MFMA) 1e-6 on the force, over 400 steps against
the oracle and over the full 25,133-step horizon (200 optical-period
equivalents at kv up to 0.5) against the FP64 path; the long-time limit
equals the steady state of molmot/obe/obe_subdoppler.py where that is
unique and reached (two-level closed form; synthetic 2+2 system)."""
import numpy as np
import pytest

from oracle import lindblad as lb
from molmot_amd.molecules import synthetic_molecule
from molmot_amd.molecules.sroh import MolecularData

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda", 0)


def sroh_pairs(mol):
    G = mol.Gamma
    return [{"freq": mol.omega_J32 + (-1.0 + 0.7) * G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0},
            {"freq": mol.omega_J32 + (-1.0 - 0.7) * G / (2 * np.pi), "q_fwd": 0, "q_bwd": 2, "s0": 1.0},
            {"freq": mol.omega_J12 + (-1.0 + 0.7) * G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0},
            {"freq": mol.omega_J12 + (-1.0 - 0.7) * G / (2 * np.pi), "q_fwd": 0, "q_bwd": 2, "s0": 1.0}]


def run_gpu(mol, pairs, v, phi, z_macro, Bp, Om, t_total, dt, precision, avg_from=0.5, pol=None):
    import torch
    from molmot_amd.obe.lindblad import LindbladClass, lindblad_rk4_device
    cls = LindbladClass(mol, pairs, z_macro, Bp, Om, 1.0, pol)
    n_steps = int(round(t_total / dt))
    d = torch.device("cuda", 0)
    F, R, rho = lindblad_rk4_device(cls, torch.tensor(np.asarray(v) * cls.kfac, device=d),
                                    torch.tensor(np.asarray(phi, dtype=np.float64), device=d), dt, n_steps,
                                    int(avg_from * n_steps), precision, want_rho=True)
    torch.cuda.synchronize()
    return F.cpu().numpy(), R.cpu().numpy(), rho.cpu().numpy()


def run_oracle(mol, pairs, v, phi, z_macro, Bp, Om, t_total, dt, avg_from=0.5, pol=None):
    out = [lb.rk4_force(mol, pairs, vi, pi, z_macro, Bp, Om, t_total, dt, avg_from=avg_from, pol=pol)
           for vi, pi in zip(v, phi)]
    return np.array([o[0] for o in out]), np.array([o[1] for o in out]), np.stack([o[2] for o in out])


def random_pol(n_pairs, seed=7):
    rng = np.random.default_rng(seed)
    pol = rng.normal(size=(n_pairs, 2, 3)) + 1j * rng.normal(size=(n_pairs, 2, 3))
    return pol / np.linalg.norm(pol, axis=-1, keepdims=True)


@pytest.mark.parametrize("pol_kind", ["integer", "random"])
@pytest.mark.parametrize("precision,tol_F,tol_rho", [("fp64", 1e-10, 1e-12), ("fp32", 1e-6, 2e-6),
                                                      ("fp32_mfma", 1e-6, 2e-6)])
def test_sroh_rk4_vs_oracle(mol, dev, precision, tol_F, tol_rho, pol_kind):
    """SrOH 12+4, four beam pairs, Omega = 1 Gamma, z = 1 mm in 0.2 T/m,
    8 (v, phi0) instances, 400 RK4 steps of 0.05/Gamma; the beams' integer
    polarisations of the reference, or seeded random complex polarisation
    vectors (the class the uneven benchmark batch uses)."""
    pairs = sroh_pairs(mol)
    pol = None if pol_kind == "integer" else random_pol(len(pairs))
    v = np.array([0.0, 0.3, -0.7, 1.5, -2.0, 0.05, 1.0, -1.2])
    phi = np.linspace(0, 2 * np.pi, 8, endpoint=False)
    Fg, Rg, rg = run_gpu(mol, pairs, v, phi, 1e-3, 0.2, 1.0, 20.0, 0.05, precision, pol=pol)
    Fo, Ro, ro = run_oracle(mol, pairs, v, phi, 1e-3, 0.2, 1.0, 20.0, 0.05, pol=pol)
    assert np.max(np.abs(Fg - Fo)) / np.max(np.abs(Fo)) < tol_F
    assert np.max(np.abs(Rg - Ro)) / np.max(np.abs(Ro)) < tol_F
    assert np.max(np.abs(rg - ro)) < tol_rho
    tr = np.trace(rg, axis1=1, axis2=2)
    assert np.max(np.abs(tr - 1)) < (1e-12 if precision == "fp64" else 1e-5)
    assert np.max(np.abs(rg - np.conj(np.transpose(rg, (0, 2, 1))))) < (1e-12 if precision == "fp64" else 1e-6)


def forty_level():
    m = synthetic_molecule(36, 4, seed=11)
    G = m.Gamma
    Ee = np.mean(m.energies[36:])
    pairs = [{"freq": Ee - np.mean(m.energies[:12]) - 1.0 * G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0},
             {"freq": Ee - np.mean(m.energies[12:36]) - 1.0 * G / (2 * np.pi), "q_fwd": 0, "q_bwd": 2, "s0": 1.0}]
    rng = np.random.default_rng(3)
    pol = rng.normal(size=(2, 2, 3)) + 1j * rng.normal(size=(2, 2, 3))
    pol /= np.linalg.norm(pol, axis=-1, keepdims=True)
    return m, pairs, pol


@pytest.mark.parametrize("pol_kind", ["integer", "random"])
@pytest.mark.parametrize("precision", ["fp32", "fp32_mfma"])
def test_fp32_long_horizon_vs_fp64(mol, dev, precision, pol_kind):
    """The full run of the benchmark's Lindblad workload, 25,133 RK4 steps of
    0.05/Gamma (force averaged over the second half), on 16 velocities in
    [-2, 2] m/s x 16 standing-wave phases: the FP32 paths against the FP64
    path (itself pinned to the oracle above), per instance and per velocity
    point (force averaged over the phases), normalised by the largest
    |F|: <= 1e-6."""
    import torch
    from molmot_amd.obe.lindblad import LindbladClass, lindblad_rk4_device
    pairs = sroh_pairs(mol)
    pol = None if pol_kind == "integer" else random_pol(len(pairs))
    cls = LindbladClass(mol, pairs, 1e-3, 0.2, 1.0, 1.0, pol)
    V, P = np.meshgrid(np.linspace(-2.0, 2.0, 16), np.linspace(0, 2 * np.pi, 16, endpoint=False), indexing="ij")
    kv = torch.from_numpy((V * cls.kfac).ravel()).to(dev)
    ph = torch.from_numpy(P.ravel()).to(dev)
    n = 25133
    out = {}
    for prec in ("fp64", precision):
        print(f"lindblad long horizon: 256 instances x {n} steps, {prec}, pol {pol_kind}", flush=True)
        F, _, _ = lindblad_rk4_device(cls, kv, ph, 0.05, n, n // 2, prec)
        torch.cuda.synchronize()
        out[prec] = F.cpu().numpy().reshape(16, 16)
    F64, F32 = out["fp64"], out[precision]
    inst = np.max(np.abs(F32 - F64)) / np.max(np.abs(F64))
    point = np.max(np.abs(F32.mean(1) - F64.mean(1))) / np.max(np.abs(F64.mean(1)))
    print(f"per instance {inst:.2e}, per point {point:.2e}", flush=True)
    assert inst < 1e-6 and point < 1e-6


@pytest.mark.parametrize("precision,tol_F", [("fp64", 1e-10), ("fp32", 1e-6)])
def test_forty_level_vs_oracle(dev, precision, tol_F):
    """Seeded synthetic 36+4 system with random complex polarisation vectors
    (the uneven-batch class), 4 instances x 100 steps."""
    m, pairs, pol = forty_level()
    v = np.array([0.0, 0.5, -1.0, 2.0])
    phi = np.array([0.0, 1.0, 2.0, 3.0])
    Fg, _, rg = run_gpu(m, pairs, v, phi, 1e-3, 0.2, 1.0, 5.0, 0.05, precision, pol=pol)
    Fo, _, ro = run_oracle(m, pairs, v, phi, 1e-3, 0.2, 1.0, 5.0, 0.05, pol=pol)
    assert np.max(np.abs(Fg - Fo)) / np.max(np.abs(Fo)) < tol_F
    assert np.max(np.abs(rg - ro)) < (1e-12 if precision == "fp64" else 2e-6)


def test_two_level_steady_state(dev):
    """One travelling wave on a two-level atom: the long-time RK4 force is
    hbar k Gamma rho_ee with the closed-form rho_ee."""
    tdm = np.zeros((2, 2, 3), dtype=complex)
    tdm[0, 1, 2] = tdm[1, 0, 2] = 1.0
    Z = np.zeros((2, 2), dtype=complex)
    m = MolecularData(energies=np.array([0.0, 4.0e14]), tdm=tdm, zeeman_x=Z, zeeman_y=Z, zeeman_z=Z,
                      n_ground=1, n_excited=1, n_states=2)
    G = m.Gamma
    pairs = [{"freq": m.energies[1] - 0.7 * G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0}]
    F, R, _ = run_gpu(m, pairs, [0.0], [0.3], 0.0, 0.0, 1.2, 60.0, 0.02, "fp64")
    Om, d = 1.2 * G, -0.7 * G
    pe = (Om ** 2 / 4) / (d ** 2 + G ** 2 / 4 + Om ** 2 / 2)
    assert abs(F[0] - lb.HBAR * m.k * G * pe) < 1e-9 * lb.HBAR * m.k * G * pe
    assert abs(R[0] - G * pe) < 1e-9 * G * pe


def test_multilevel_long_time_is_steady_state(dev):
    """Synthetic 2+2 system (unique steady state, gap ~0.16 Gamma): GPU FP64
    RK4 over 400/Gamma equals the oracle's lstsq steady state."""
    m = synthetic_molecule(2, 2, seed=4)
    G = m.Gamma
    Ee = np.mean(m.energies[2:])
    pairs = [{"freq": Ee - m.energies[0] - 0.8 * G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0},
             {"freq": Ee - m.energies[1] - 0.5 * G / (2 * np.pi), "q_fwd": 1, "q_bwd": 1, "s0": 1.0}]
    phi = 0.4
    H = lb.hamiltonian(m, pairs, 0.0, 2e-3, phi, 0.5, 1.0)
    rho_ss = lb.steady_state(lb.liouvillian(H, m, G), m.n_states)
    A, B = lb.coupling_blocks(m, pairs, 1.0)
    F_ss = lb.force_from_rho(rho_ss, m, A * np.exp(1j * phi), B * np.exp(-1j * phi))
    F, _, rho = run_gpu(m, pairs, [0.0], [phi], 2e-3, 0.5, 1.0, 400.0, 0.02, "fp64", avg_from=0.9)
    assert np.max(np.abs(rho[0] - rho_ss)) < 1e-10
    assert abs(F[0] - F_ss) < 1e-9 * abs(F_ss)


def _augmented(L, N):
    M = L.copy()
    M[-1, :] = 0
    M[-1, [i * N + i for i in range(N)]] = 1.0
    return M


@pytest.mark.parametrize("phi", [0.0, 0.7, 4.0])
def test_liouvillian_kernel_vs_oracle(mol, dev, phi):
    """Augmented Liouvillian from the kernel equals the oracle's
    build_liouvillian_fast restatement / Gamma with the trace row."""
    import torch
    from molmot_amd.obe.lindblad import LindbladClass, liouvillian_device
    pairs = sroh_pairs(mol)
    cls = LindbladClass(mol, pairs, 1e-3, 0.2, 1.0)
    ph = torch.tensor([phi, 0.1], dtype=torch.float64, device=dev)
    M = liouvillian_device(cls, ph).cpu().numpy()
    L = liouvillian_device(cls, ph, trace_row=False).cpu().numpy()
    N = mol.n_states
    H = lb.hamiltonian(mol, pairs, 0.0, 1e-3, phi, 0.2, 1.0)
    ref = lb.liouvillian(H, mol, mol.Gamma) / mol.Gamma
    assert M.shape == (2, N * N, N * N)
    assert np.max(np.abs(M[0] - _augmented(ref, N))) < 1e-13 * np.max(np.abs(ref))
    assert np.max(np.abs(L[0] - ref)) < 1e-13 * np.max(np.abs(ref))


def test_steady_state_two_level_closed_form(dev):
    from molmot_amd.obe.lindblad import obe_force_subdoppler
    tdm = np.zeros((2, 2, 3), dtype=complex)
    tdm[0, 1, 2] = tdm[1, 0, 2] = 1.0
    Z = np.zeros((2, 2), dtype=complex)
    m = MolecularData(energies=np.array([0.0, 4.0e14]), tdm=tdm, zeeman_x=Z, zeeman_y=Z, zeeman_z=Z,
                      n_ground=1, n_excited=1, n_states=2)
    G = m.Gamma
    for det, Om_G in [(0.0, 0.5), (-1.0, 1.0), (2.0, 3.0)]:
        pairs = [{"freq": m.energies[1] + det * G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0}]
        F, rho, R = obe_force_subdoppler(m, pairs, 0.0, 0.0, 0.0, Om_G, n_spatial=4)
        Om, d = Om_G * G, det * G
        pe = (Om ** 2 / 4) / (d ** 2 + G ** 2 / 4 + Om ** 2 / 2)
        assert abs(rho[1, 1].real - pe) < 1e-12
        assert abs(F - lb.HBAR * m.k * G * pe) < 1e-10 * lb.HBAR * m.k * G
        assert abs(R - G * pe) < 1e-9 * G


@pytest.mark.parametrize("Bp,z", [(0.2, 1e-3), (1.0, 2e-3)])
def test_subdoppler_force_sroh_vs_oracle(mol, dev, Bp, z):
    """SrOH, four beam pairs, B != 0: GPU steady-state force, rho_avg and R vs
    the oracle's lstsq restatement of obe_force_subdoppler over 16 phases.
    14 phases have a unique steady state (condition number ~1e4); at
    phi = pi/2 and 3 pi/2 the forward and backward couplings cancel, every
    ground-state operator is stationary and the minimum-norm solution (uniform
    ground populations) is the reference's answer."""
    from molmot_amd.obe.lindblad import obe_force_subdoppler, obe_force_subdoppler_hip
    pairs = sroh_pairs(mol)
    Fo, ro, Ro = lb.force_subdoppler(mol, pairs, 0.0, z, Bp, 1.0, n_spatial=16)
    F, rho, R = obe_force_subdoppler(mol, pairs, 0.0, z, Bp, 1.0, n_spatial=16)
    assert abs(F - Fo) < 1e-9 * max(abs(Fo), 1e-3 * lb.HBAR * mol.k * mol.Gamma)
    assert np.max(np.abs(rho - ro)) < 1e-10
    assert abs(R - Ro) < 1e-9 * Ro
    Fv, rv, Rv = obe_force_subdoppler_hip(mol, pairs, np.array([0.0, 1.0, 2.0]), z, Bp, 1.0, n_spatial=16)
    assert Fv.shape == (3,) and rv.shape == (3, mol.n_states, mol.n_states) and Rv.shape == (3,)
    assert np.max(np.abs(Fv - Fo)) < 1e-8 * max(abs(Fo), 1e-3 * lb.HBAR * mol.k * mol.Gamma)


def test_subdoppler_forty_level_vs_oracle(dev):
    """36+4 synthetic class with complex polarisation vectors (N^2 = 1600
    augmented system, condition number ~2e7) at one phase: kernel
    Liouvillian elementwise, the GPU steady state's residual against the
    oracle's Liouvillian, and agreement with the oracle's steady state to the
    conditioning bound."""
    import torch
    from molmot_amd.obe.lindblad import LindbladClass, liouvillian_device, steady_state_device
    m = synthetic_molecule(36, 4, seed=11)
    G = m.Gamma
    N = m.n_states
    Ee = np.mean(m.energies[36:])
    pairs = [{"freq": Ee - np.mean(m.energies[:12]) - 1.0 * G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0},
             {"freq": Ee - np.mean(m.energies[12:36]) - 1.0 * G / (2 * np.pi), "q_fwd": 0, "q_bwd": 2, "s0": 1.0}]
    rng = np.random.default_rng(3)
    pol = rng.normal(size=(2, 2, 3)) + 1j * rng.normal(size=(2, 2, 3))
    pol /= np.linalg.norm(pol, axis=-1, keepdims=True)
    phi = 0.9
    cls = LindbladClass(m, pairs, 1e-3, 0.2, 1.0, 1.0, pol)
    ph = torch.tensor([phi], dtype=torch.float64, device=dev)
    M = liouvillian_device(cls, ph).cpu().numpy()[0]
    L = lb.liouvillian(lb.hamiltonian(m, pairs, 0.0, 1e-3, phi, 0.2, 1.0, pol), m, G) / G
    assert np.max(np.abs(M - _augmented(L, N))) < 1e-13 * np.max(np.abs(L))
    rho = steady_state_device(cls, ph, clip_negative=False).cpu().numpy()[0]
    x = rho.reshape(-1, order="F")
    assert np.max(np.abs(L @ x)) < 1e-11 * np.max(np.abs(L))
    assert abs(np.trace(rho) - 1) < 1e-13
    ro = lb.steady_state(L * G, N)
    assert np.max(np.abs(rho - ro)) < 1e-7


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lanes_per_instance_agree(mol, dev, precision):
    """64, 128 and 256 lanes per instance run the same per-element arithmetic;
    only the order of the final force/rate sums differs (1e-12 relative in
    FP64, 1e-6 in FP32), and rho agrees to the last bits."""
    import torch
    from molmot_amd.obe.lindblad import LindbladClass, lindblad_rk4_device
    m40, pairs40, pol40 = forty_level()
    for m, pairs, pol in ((mol, sroh_pairs(mol), None), (m40, pairs40, pol40)):
        cls = LindbladClass(m, pairs, 1e-3, 0.2, 1.0, 1.0, pol)
        kv = torch.tensor(np.array([0.0, 0.4, -1.3, 2.0, 0.9]) * cls.kfac, device=dev)
        ph = torch.tensor(np.array([0.1, 2.0, 4.0, 5.5, 3.0]), device=dev)
        res = {}
        for lanes in (64, 128, 256):
            cls.set_lanes(precision, lanes)
            assert cls.lanes(precision) == lanes
            F, R, rho = lindblad_rk4_device(cls, kv, ph, 0.05, 60, 30, precision, want_rho=True)
            res[lanes] = (F.cpu().numpy(), R.cpu().numpy(), rho.cpu().numpy())
        cls.set_lanes(precision, 0)
        assert cls.lanes(precision) == (64 if cls.N == 16 else 256)
        tol = 1e-12 if precision == "fp64" else 1e-6
        for lanes in (128, 256):
            F, R, rho = res[lanes]
            assert np.max(np.abs(F - res[64][0])) <= tol * np.max(np.abs(res[64][0]))
            assert np.max(np.abs(R - res[64][1])) <= tol * np.max(np.abs(res[64][1]))
            assert np.max(np.abs(rho - res[64][2])) <= (1e-15 if precision == "fp64" else 1e-7)


@pytest.mark.parametrize("precision", ["fp64", "fp32_mfma"])
def test_uneven_launch_equals_separate_launches(mol, dev, precision):
    """16- and 40-level instances in one launch (five 16-level instances: a
    block with three spare wavefronts) give the same F, R, rho as the two
    classes launched separately with their default lanes."""
    import torch
    from molmot_amd.obe.lindblad import LindbladClass, lindblad_rk4_device, lindblad_rk4_uneven_device
    m40, pairs40, pol40 = forty_level()
    c16 = LindbladClass(mol, sroh_pairs(mol), 1e-3, 0.2, 1.0, 1.0, None)
    c40 = LindbladClass(m40, pairs40, 1e-3, 0.2, 1.0, 1.0, pol40)
    kv16 = torch.tensor(np.array([0.0, 0.4, -1.3, 2.0, 0.9]) * c16.kfac, device=dev)
    ph16 = torch.tensor(np.array([0.1, 2.0, 4.0, 5.5, 3.0]), device=dev)
    kv40 = torch.tensor(np.array([0.7, -0.2, 1.1]) * c40.kfac, device=dev)
    ph40 = torch.tensor(np.array([1.0, 2.5, 6.0]), device=dev)
    (F16, R16, r16), (F40, R40, r40) = lindblad_rk4_uneven_device(c16, kv16, ph16, c40, kv40, ph40, 0.05, 60, 30,
                                                                  precision, want_rho=True)
    p40 = "fp64" if precision == "fp64" else "fp32"
    a16 = lindblad_rk4_device(c16, kv16, ph16, 0.05, 60, 30, precision, want_rho=True)
    a40 = lindblad_rk4_device(c40, kv40, ph40, 0.05, 60, 30, p40, want_rho=True)
    for got, ref in ((F16, a16[0]), (R16, a16[1]), (r16, a16[2]), (F40, a40[0]), (R40, a40[1]), (r40, a40[2])):
        assert torch.equal(got, ref)
