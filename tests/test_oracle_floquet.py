"""Pin the Floquet oracle (oracle/floquet.py): the reference's own FloquetOBE
outputs (tests/golden/ref_floquet.npz, molmot/obe/floquet.py:61-403), and for
the standing-wave model the adiabatic limit: at v -> 0 the time-averaged
force is the average over the standing-wave phase of the steady-state force
of oracle.lindblad (itself pinned to obe_subdoppler vectors)."""
import os

import numpy as np
import pytest

from oracle import floquet as fq
from oracle import lindblad as lb
from molmot_amd.molecules.sroh import MolecularData

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLDEN, "ref_floquet.npz"), allow_pickle=False)


def _syn(d):
    return MolecularData(energies=d["syn_energies"], tdm=d["syn_tdm"], zeeman_x=d["syn_zeeman_x"],
                         zeeman_y=d["syn_zeeman_y"], zeeman_z=d["syn_zeeman_z"], n_ground=int(d["syn_n_ground"]),
                         n_excited=int(d["syn_n_excited"]), n_states=6)


def _pairs(d, tag):
    return [dict(freq=float(f), q_fwd=int(a), q_bwd=int(b), s0=float(s))
            for f, a, b, s in zip(d[tag + "_pairs_freq"], d[tag + "_pairs_q_fwd"], d[tag + "_pairs_q_bwd"],
                                  d[tag + "_pairs_s0"])]


@pytest.mark.parametrize("tag", ["sroh", "syn"])
def test_floquet_vs_reference(ref, mol, tag):
    m = mol if tag == "sroh" else _syn(ref)
    nmax, Bp, zm, Om, ge = ref[tag + "_args"]
    nmax = int(nmax)
    blk = fq.blocks(m, _pairs(ref, tag), Om, ge)
    v = ref[tag + "_v"]
    scale = fq.HBAR * m.k * m.Gamma
    F = fq.force(m, blk, v, nmax)
    assert np.max(np.abs(F - ref[tag + "_F"])) < 1e-10 * scale
    Fz = fq.force(m, blk, v, nmax, Bp, zm)
    assert np.max(np.abs(Fz - ref[tag + "_F_z"])) < 1e-10 * scale
    R = np.array([fq.scattering_rate(m, blk, x, nmax, ge) for x in v])
    assert np.allclose(R, ref[tag + "_R"], rtol=1e-10, atol=1e-12 * m.Gamma)   # SrOH at z = 0: dark states, R ~ 1e-12 Gamma
    for i, x in enumerate(v):
        rho0 = fq.solve(m, blk, x if abs(x) >= 1e-12 else 1e-6, nmax, Bp, zm)[nmax]
        assert np.max(np.abs(rho0 - ref[tag + "_rho0"][i])) < 1e-10
    if tag == "syn":
        assert np.max(np.abs(ref["syn_F"])) > 1e-3 * scale      # a force well above rounding


@pytest.mark.parametrize("x,periods", [(0.25, 12), (0.6, 40)])
def test_standing_wave_matches_time_domain(ref, x, periods):
    """Moving atom, kv = x Gamma: the Floquet force of the standing-wave
    model (n_max = 12) equals the time average of the RK4 propagation of
    oracle.lindblad.rk4_force (400 steps per period of the standing-wave
    phase, averaged over whole periods after the transients): 2e-6."""
    m = _syn(ref)
    pairs = _pairs(ref, "syn")
    blk = fq.blocks(m, pairs, 0.8, 1.0, coupling="standing_wave")
    v = x * m.Gamma / m.k
    T = 2 * np.pi / x
    n = periods * 400
    F_rk4, _, _ = lb.rk4_force(m, pairs, v, 0.0, 0.0, 0.0, 0.8, periods * T, T / 400, avg_from=(n // 2 + 0.5) / n)
    F = fq.force(m, blk, [v], 12)[0]
    assert abs(F - F_rk4) < 2e-6 * abs(F)
    # a wrong sideband (rho_+1 with A) or a dropped backward term fails
    rc = fq.solve(m, blk, v, 12)
    wrong = 2 * fq.HBAR * m.k * (np.imag(np.trace(rc[13] @ blk["A"])) - np.imag(np.trace(rc[11] @ blk["B"])))
    assert abs(wrong - F_rk4) > 1e-2 * abs(F)


def test_standing_wave_adiabatic_limit(ref):
    """v = 1e-6 m/s: the force tends to the average over the standing-wave
    phase of the steady-state force (64 phases); the Fourier series of the
    adiabatically followed state converges slowly, so n_max = 14 gives it to
    ~1%."""
    m = _syn(ref)
    pairs = _pairs(ref, "syn")
    blk = fq.blocks(m, pairs, 0.8, 1.0, coupling="standing_wave")
    F = fq.force(m, blk, [1e-6], 14)[0]
    A, B = lb.coupling_blocks(m, pairs, 0.8)
    Fs = []
    for p in np.linspace(0, 2 * np.pi, 64, endpoint=False):
        H = lb.hamiltonian(m, pairs, 0.0, 0.0, p, 0.0, 0.8)
        rho = lb.steady_state(lb.liouvillian(H, m, m.Gamma), 6)
        Fs.append(lb.force_from_rho(rho, m, A * np.exp(1j * p), B * np.exp(-1j * p)))
    assert abs(F - np.mean(Fs)) < 1e-2 * abs(np.mean(Fs))
    assert abs(np.mean(Fs)) > 1e-2 * np.max(np.abs(Fs))
