"""Write small input/output vectors of reference functions that no bundled
fixture, closed form or invariant pins on its own.

Run in the development container only (it imports the reference package from
/root/reference, which never travels with this repository):

    python tests/golden/make_reference_fixtures.py

Writes, next to this script:

* ref_subdoppler.npz -- molmot/obe/obe_subdoppler.py:19-286:
  build_H_standing_wave at several standing-wave phases and
  obe_force_subdoppler (F, rho_avg, R) for SrOH with the 4-frequency DC pairs
  at two (z, B') with B' != 0, SrOH with the RF and blue pairs, and a seeded
  4+2 system with complex dipoles and a non-zero excited Zeeman diagonal.
* ref_force3d.npz -- molmot/mot/simulator_3d.py:59-365: rate_eq_force_3d on
  make_6beam_config beams for a seeded non-degenerate 12+4 molecule at 32
  (r, v) points, and for SrOH at 8 points.
* ref_force3d_axis.npz -- rate_eq_force_3d for SrOH at 8 points on the +z
  axis (B along +z), where the GPU's spin-rotation basis is the eigh basis.
* ref_traj3d_mc.npz -- molmot/propagation/trajectories.py:78-229:
  simulate_trajectories_3d (4 particles, 400 steps, one escapes) and
  monte_carlo_temperature (4 particles, 300 steps) on DCMOTSimulator
  (molmot/mot/simulator.py:93-162), NumPy rate equations (numba is absent).
* ref_analysis.npz -- molmot/mot/analysis.py:128-288: sigma_vs_time,
  temperature_vs_time and density_vs_time of 24 seeded random-walk
  (t, z, v) lists of uneven length, some escape-filled beyond 15 mm.
* ref_floquet.npz -- molmot/obe/floquet.py:61-403: FloquetOBE force,
  scattering rate and rho_0 for SrOH with the 4-frequency DC pairs (n_max
  2, three velocities, z = 0 and z != 0) and for the seeded 4+2 system
  (n_max 3, five velocities including v = 0).
* ref_tables.npz -- molmot/molecules/{sroh,caoh,caf}.py and molmot/states/:
  the tables of build_sroh_hamiltonian, build_caoh_hamiltonian,
  build_caf_hamiltonian, load_caoh_from_sroh_structure and
  load_caf_from_sroh_structure; for each molecule the ground and excited
  Hamiltonian matrices, their eigenvectors, the ground Zeeman and nuclear
  Zeeman matrices, the basis-level dipole matrix and the case (a) -> (b)
  overlap matrix; and the case (b) / case (a) operators on larger bases
  (N = 0..2; J = 1/2, 3/2 with |P| = 1/2, 3/2).
* ref_sse.npz -- molmot/obe/stochastic.py:103-277, 321-556, 806-954:
  rotate_polarization for on- and off-plane directions, compute_fields_6beam
  and get_omega0s_zeeman of an SSEProblem, and SSESolver._rhs at fixed
  (t, psi, r, v).

Only inputs and outputs are stored; molecular tables of the synthetic systems
are stored so that the tests rebuild them without this script.
"""
import os
import sys

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def _import_reference():
    sys.path.insert(0, REF)
    from molmot.molecules.sroh import MolecularData, load_sroh_from_julia
    from molmot.obe import obe_subdoppler as osd
    from molmot.mot import simulator_3d as s3d
    from molmot.mot.simulator import DCMOTSimulator
    from molmot.propagation import trajectories as trj
    from molmot.obe import stochastic as sto
    return MolecularData, load_sroh_from_julia, osd, s3d, DCMOTSimulator, trj, sto


def synthetic_tables(ng, ne, seed, excited_zeeman=False, nondegenerate=True):
    """Seeded complex dipoles (normalised per excited state) and Hermitian
    Zeeman blocks; energies in two ground manifolds ~55 MHz apart."""
    rng = np.random.default_rng(seed)
    ns = ng + ne
    E = np.empty(ns)
    nlo = max(ng // 3, 1)
    E[:nlo] = 1.4868e10 + rng.uniform(0.0, 2e5, nlo)
    E[nlo:ng] = 1.4977e10 + rng.uniform(0.0, 2e6, ng - nlo)
    E[ng:] = 4.320351078685233e14 + rng.uniform(-3e6, 3e6, ne)
    tdm = np.zeros((ns, ns, 3), dtype=complex)
    m_g = np.array([(i % 3) - 1 for i in range(ng)])
    m_e = np.array([(i % 3) - 1 for i in range(ne)])
    for ie in range(ne):
        cols = []
        for ig in range(ng):
            q = m_e[ie] - m_g[ig]
            if abs(q) <= 1:
                tdm[ig, ng + ie, q + 1] = rng.normal() + 1j * rng.normal()
                cols.append((ig, q + 1))
        nrm = np.sqrt(sum(abs(tdm[ig, ng + ie, qi]) ** 2 for ig, qi in cols))
        for ig, qi in cols:
            tdm[ig, ng + ie, qi] /= nrm
            tdm[ng + ie, ig, qi] = np.conj(tdm[ig, ng + ie, qi])
    scale = 0.5 * 1.39962449361e6 / 1.0186e6
    zz = np.zeros((ns, ns), dtype=complex)
    zz[:ng, :ng] = np.diag(scale * (m_g + (0.37 * rng.uniform(-1, 1, ng) if nondegenerate else 0.0)))
    off = 0.02 * scale * (rng.normal(size=(ng, ng)) + 1j * rng.normal(size=(ng, ng)))
    off = np.triu(off, 1)
    zz[:ng, :ng] += off + off.conj().T
    if excited_zeeman:
        zz[ng:, ng:] = np.diag(0.3 * scale * rng.uniform(-1, 1, ne))
    zx = np.zeros((ns, ns), dtype=complex)
    zy = np.zeros((ns, ns), dtype=complex)
    for ig in range(ng - 1):
        a = scale * (0.5 + 0.1 * rng.uniform())
        zx[ig, ig + 1] = zx[ig + 1, ig] = a
        zy[ig, ig + 1] = -1j * a
        zy[ig + 1, ig] = 1j * a
    return dict(energies=E, tdm=tdm, zeeman_x=zx, zeeman_y=zy, zeeman_z=zz, n_ground=ng, n_excited=ne)


def _pairs_arrays(pairs):
    return dict(freq=np.array([p["freq"] for p in pairs]), q_fwd=np.array([p["q_fwd"] for p in pairs]),
                q_bwd=np.array([p["q_bwd"] for p in pairs]), s0=np.array([p["s0"] for p in pairs]))


def subdoppler(MolecularData, sroh, osd):
    out = {}
    cases = []
    dc4 = osd.make_beam_pairs_dc4(sroh, -1.0, 0.5, 2.0)
    for (zm, Bp, v) in ((1.2e-3, 0.21, 0.3), (-2.5e-3, 0.15, -1.0)):
        cases.append(("sroh_dc4", sroh, dc4, v, zm, Bp, 1.0, 16))
    cases.append(("sroh_rf", sroh, osd.make_beam_pairs_rf(sroh, -0.8, 1.5, phase=1), 0.0, 0.8e-3, 0.2, 0.7, 8))
    cases.append(("sroh_blue", sroh, osd.make_beam_pairs_blue(sroh, 2.0, -0.4, 1.2), 0.0, -0.6e-3, 0.3, 1.3, 8))
    syn = synthetic_tables(4, 2, seed=21, excited_zeeman=True)
    msyn = MolecularData(**syn, n_states=6)
    Gh = msyn.Gamma / (2 * np.pi)
    Ee = np.mean(msyn.energies[4:])
    syn_pairs = [dict(freq=Ee - np.mean(msyn.energies[:1]) - 0.7 * Gh, q_fwd=2, q_bwd=0, s0=1.5),
                 dict(freq=Ee - np.mean(msyn.energies[1:4]) + 0.4 * Gh, q_fwd=0, q_bwd=1, s0=0.8),
                 dict(freq=Ee - np.mean(msyn.energies[1:4]) - 1.3 * Gh, q_fwd=1, q_bwd=2, s0=2.5)]
    for k, val in syn.items():
        out["syn_" + k] = val
    cases.append(("syn42", msyn, syn_pairs, 0.5, 2.0e-3, 0.25, 0.9, 8))
    names = []
    for name, m, pairs, v, zm, Bp, Om, ns in cases:
        i = len(names)
        names.append(name)
        for k, val in _pairs_arrays(pairs).items():
            out[f"c{i}_pairs_{k}"] = val
        out[f"c{i}_args"] = np.array([v, zm, Bp, Om, ns], dtype=float)
        phases = np.array([0.0, 0.7, 2.1, np.pi / 2])
        out[f"c{i}_phases"] = phases
        out[f"c{i}_H"] = np.stack([osd.build_H_standing_wave(m, pairs, v, zm, p, Bp, Om) for p in phases])
        F, rho, R = osd.obe_force_subdoppler(m, pairs, v, zm, Bp, Om, n_spatial=ns, Gamma_eff_factor=0.9)
        out[f"c{i}_F"] = np.array(F)
        out[f"c{i}_rho"] = rho
        out[f"c{i}_R"] = np.array(R)
        # the phase-averaged force of a standing wave nearly cancels, so the
        # force of every phase is kept too (the same reference steps as
        # obe_force_subdoppler, one phase at a time)
        meff = type(m).__new__(type(m))
        meff.__dict__.update(m.__dict__)
        meff.Gamma = m.Gamma * 0.9
        Fph = []
        for p in np.linspace(0, 2 * np.pi, ns, endpoint=False):
            H = osd.build_H_standing_wave(m, pairs, v, zm, p, Bp, Om)
            r = osd.solve_steady_state(osd.build_liouvillian_fast(H, meff), m.n_states)
            Fph.append(osd.force_from_rho(r, m, pairs, p, Om))
        out[f"c{i}_F_phase"] = np.array(Fph)
        print(f"subdoppler {name}: F = {F:.6e} N, R = {R:.6e}")
    out["names"] = np.array(names)
    out["gamma_eff_factor"] = np.array(0.9)
    np.savez_compressed(os.path.join(HERE, "ref_subdoppler.npz"), **out)


def force3d(MolecularData, sroh, osd, s3d):
    out = {}
    rng = np.random.default_rng(7)
    syn = synthetic_tables(12, 4, seed=33)
    msyn = MolecularData(**syn, n_states=16)
    for k, val in syn.items():
        out["syn_" + k] = val
    for tag, m, npts in (("syn", msyn, 32), ("sroh", sroh, 8)):
        pairs = osd.make_beam_pairs_dc4(m, -0.9, 0.6, 1.7)
        beams = s3d.make_6beam_config(m, pairs, beam_radius=5e-3)
        r = rng.uniform(-4e-3, 4e-3, (npts, 3))
        v = rng.uniform(-3.0, 3.0, (npts, 3))
        r[0] = 0.0
        v[0] = 0.0
        r[1] = [0.0, 0.0, 1.5e-3]
        r[2] = [2e-3, 0.0, 0.0]
        F = np.zeros((npts, 3))
        R = np.zeros(npts)
        P = np.zeros((npts, m.n_ground))
        for i in range(npts):
            F[i], R[i], P[i] = s3d.rate_eq_force_3d(m, beams, r[i], v[i], 0.18, Gamma_eff_factor=0.8)
        for k, val in _pairs_arrays(pairs).items():
            out[f"{tag}_pairs_{k}"] = val
        out[f"{tag}_r"] = r
        out[f"{tag}_v"] = v
        out[f"{tag}_F"] = F
        out[f"{tag}_R"] = R
        out[f"{tag}_P"] = P
        print(f"force3d {tag}: max|F| = {np.max(np.abs(F)):.3e} N")
    out["B_gradient"] = np.array(0.18)
    out["gamma_eff_factor"] = np.array(0.8)
    out["beam_radius"] = np.array(5e-3)
    np.savez_compressed(os.path.join(HERE, "ref_force3d.npz"), **out)


def force3d_axis(sroh, osd, s3d):
    """rate_eq_force_3d for SrOH at points on the +z axis (B along +z)."""
    rng = np.random.default_rng(19)
    pairs = osd.make_beam_pairs_dc4(sroh, -0.9, 0.6, 1.7)
    beams = s3d.make_6beam_config(sroh, pairs, beam_radius=5e-3)
    z = np.array([0.2e-3, 0.7e-3, 1.5e-3, 2.4e-3, 3.3e-3, 4.1e-3, 5.0e-3, 6.2e-3])
    r = np.c_[np.zeros((8, 2)), z]
    v = rng.uniform(-3.0, 3.0, (8, 3))
    F = np.zeros((8, 3))
    R = np.zeros(8)
    P = np.zeros((8, 12))
    for i in range(8):
        F[i], R[i], P[i] = s3d.rate_eq_force_3d(sroh, beams, r[i], v[i], 0.18, Gamma_eff_factor=0.8)
    out = {f"pairs_{k}": val for k, val in _pairs_arrays(pairs).items()}
    out.update(r=r, v=v, F=F, R=R, P=P, B_gradient=np.array(0.18), gamma_eff_factor=np.array(0.8),
               beam_radius=np.array(5e-3))
    print(f"force3d_axis: max|F| = {np.max(np.abs(F)):.3e} N")
    np.savez_compressed(os.path.join(HERE, "ref_force3d_axis.npz"), **out)


def traj3d_mc(sroh, DCMOTSimulator, trj):
    sim = DCMOTSimulator(sroh, delta_Gamma=-0.2, split_Gamma=0.7, s0=1.0, B_gradient=0.16)
    r0 = [np.array([1e-3, -0.5e-3, 2e-3]), np.array([0.0, 0.0, 0.0]),
          np.array([-3e-3, 2e-3, -1e-3]), np.array([0.0, 1e-3, 0.0])]
    v0 = [np.array([0.5, -1.0, 2.0]), np.array([-3.0, 1.0, 0.5]),
          np.array([1.0, 1.0, -1.0]), np.array([0.0, 40.0, 0.0])]
    dt, n_steps = 1e-5, 400
    tr = trj.simulate_trajectories_3d(sim, r0, v0, t_max=n_steps * dt + 0.5 * dt, dt=dt)
    out = dict(r0=np.array(r0), v0=np.array(v0), dt=np.array(dt), n_steps=np.array(len(tr[0][0])),
               r_traj=np.stack([t[1] for t in tr]), v_traj=np.stack([t[2] for t in tr]),
               sim_args=np.array([-0.2, 0.7, 1.0, 0.16]))
    print("traj3d: final |r| =", [float(np.linalg.norm(t[1][-1])) for t in tr])
    mc = dict(n_particles=4, t_max=3e-3, dt=1e-5, z_init_sigma=1e-3, v_init_sigma=1.0, t_equilibrium=1e-3)
    T = trj.monte_carlo_temperature(sim, **mc)
    out["mc_args"] = np.array([mc[k] for k in ("n_particles", "t_max", "dt", "z_init_sigma",
                                               "v_init_sigma", "t_equilibrium")], dtype=float)
    out["mc_T"] = np.array(T)
    print(f"monte_carlo_temperature: T = {T:.6e} K")
    np.savez_compressed(os.path.join(HERE, "ref_traj3d_mc.npz"), **out)


def sse(sroh, osd, sto):
    out = {}
    rng = np.random.default_rng(11)
    pols = [np.array([1, 0, 0], complex), np.array([0, 1, 0], complex), np.array([0, 0, 1], complex),
            np.array([0.3 - 0.2j, 0.5 + 0.1j, -0.4 + 0.6j])]
    khats = [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
             [1, 1, 1], [0.3, -0.5, 0.8], [-0.2, 0.1, -0.97], [0.6, 0.8, 0.0], [-0.7, 0.2, 0.4]]
    rot = np.zeros((len(pols), len(khats), 3), dtype=complex)
    for i, p in enumerate(pols):
        for j, k in enumerate(khats):
            rot[i, j] = sto.rotate_polarization(p, np.array(k, float))
    out.update(rot_pols=np.array(pols), rot_khats=np.array(khats, float), rot_out=rot)
    pairs = osd.make_beam_pairs_dc4(sroh, -0.9, 0.6, 1.7)
    prob = sto.SSEProblem(sroh, pairs, 0.2, beam_radius=5e-3)
    solver = sto.SSESolver(prob)
    for k, val in _pairs_arrays(pairs).items():
        out["pairs_" + k] = val
    N = prob.n_states
    n_pts = 4
    ts = np.array([0.0, 0.37, 5.2, 41.9])
    rs = rng.uniform(-200.0, 200.0, (n_pts, 3))
    vs = rng.uniform(-0.05, 0.05, (n_pts, 3))
    psis = rng.normal(size=(n_pts, N)) + 1j * rng.normal(size=(n_pts, N))
    psis *= 0.9 / np.linalg.norm(psis, axis=1, keepdims=True)
    E = np.zeros((n_pts, 6, 3), dtype=complex)
    W = np.zeros((n_pts, N))
    du = np.zeros((n_pts, prob.state_size))
    for i in range(n_pts):
        W[i] = prob.get_omega0s_zeeman(rs[i])
        E[i] = sto.compute_fields_6beam(rs[i], ts[i], W[i], prob.omega_lasers, prob.sats, prob.pols,
                                        prob.beam_radius_k, cached_eps=solver._cached_eps)[0]
        u = np.concatenate([psis[i].real, psis[i].imag, rs[i], vs[i]])
        du[i] = solver._rhs(ts[i], u).copy()
    out.update(t=ts, r=rs, v=vs, psi=psis, fields=E, omega0s=W, du=du, eps=solver._cached_eps,
               B_gradient=np.array(0.2), beam_radius=np.array(5e-3))
    print("sse: max |du| =", float(np.max(np.abs(du))))
    np.savez_compressed(os.path.join(HERE, "ref_sse.npz"), **out)


def analysis():
    sys.path.insert(0, REF)
    from molmot.mot import analysis as an
    rng = np.random.default_rng(5)
    n = 24
    lens = rng.integers(250, 320, n)
    T = np.zeros((n, 320))
    Z = np.full((n, 320), np.nan)
    V = np.full((n, 320), np.nan)
    results = []
    for i in range(n):
        m = int(lens[i])
        t = np.arange(m) * 1e-5
        v = np.cumsum(rng.normal(0, 0.3, m)) + rng.normal(0, 4.0)
        z = np.cumsum(v) * 1e-5 + rng.normal(0, 7e-3)
        esc = np.nonzero(np.abs(z) > 0.015)[0]
        if esc.size:
            z[esc[0]:] = z[esc[0]]
            v[esc[0]:] = v[esc[0]]
        T[i, :m], Z[i, :m], V[i, :m] = t, z, v
        results.append((t, z, v))
    tc, sz, sv, na = an.sigma_vs_time(results, time_bins=20)
    _, Tk = an.temperature_vs_time(results, 1.7e-25, time_bins=20)
    _, n0 = an.density_vs_time(results, N_total=3000, time_bins=20)
    print("analysis: alive", na.min(), na.max())
    np.savez_compressed(os.path.join(HERE, "ref_analysis.npz"), lens=lens, t=T, z=Z, v=V, t_centres=tc, sigma_z=sz,
                        sigma_v=sv, n_alive=na, T=Tk, n0=n0, mass=np.array(1.7e-25), N_total=np.array(3000),
                        time_bins=np.array(20))


def floquet(MolecularData, sroh, osd):
    from molmot.obe.floquet import FloquetOBE
    out = {}
    syn = synthetic_tables(4, 2, seed=21, excited_zeeman=True)
    msyn = MolecularData(**syn, n_states=6)
    Gh = msyn.Gamma / (2 * np.pi)
    Ee = np.mean(msyn.energies[4:])
    syn_pairs = [dict(freq=Ee - np.mean(msyn.energies[:1]) - 0.7 * Gh, q_fwd=2, q_bwd=0, s0=1.5),
                 dict(freq=Ee - np.mean(msyn.energies[1:4]) + 0.4 * Gh, q_fwd=0, q_bwd=1, s0=0.8)]
    for k, val in syn.items():
        out["syn_" + k] = val
    cases = [("sroh", sroh, osd.make_beam_pairs_dc4(sroh, -1.0, 0.5, 2.0), 2, [0.3, -1.1, 2.5], 0.2, 1.5e-3),
             ("syn", msyn, syn_pairs, 3, [0.0, 0.05, -0.4, 1.0, 3.0], 0.25, -2e-3)]
    for tag, m, pairs, nmax, vs, Bp, zm in cases:
        fo = FloquetOBE(m, pairs, Bp, Omega_Gamma=0.8, n_max=nmax, Gamma_eff_factor=0.9)
        for k, val in _pairs_arrays(pairs).items():
            out[f"{tag}_pairs_{k}"] = val
        out[f"{tag}_args"] = np.array([nmax, Bp, zm, 0.8, 0.9])
        out[f"{tag}_v"] = np.array(vs)
        out[f"{tag}_F"] = fo.force_vs_velocity(np.array(vs))
        out[f"{tag}_F_z"] = fo.force_vs_velocity(np.array(vs), z_macro=zm)
        out[f"{tag}_R"] = np.array([fo.scattering_rate(v) for v in vs])
        out[f"{tag}_rho0"] = np.stack([fo.steady_state_rho(v, z_macro=zm) for v in vs])
        print(f"floquet {tag}: F = {out[tag + '_F']}")
    np.savez_compressed(os.path.join(HERE, "ref_floquet.npz"), **out)


def tables():
    sys.path.insert(0, REF)
    from fractions import Fraction
    from molmot.molecules import sroh as msr, caoh as mca, caf as mcf
    from molmot.states.basis import enumerate_states
    from molmot.states import case_a as ca, case_b as cb
    from molmot.states.hamiltonian import Hamiltonian, operator_to_matrix
    from molmot.states.overlaps import overlap_caseb_casea
    out = {}
    builders = [("sroh_build", msr.build_sroh_hamiltonian), ("caoh_build", mca.build_caoh_hamiltonian),
                ("caf_build", mcf.build_caf_hamiltonian),
                ("caoh_load", mca.load_caoh_from_sroh_structure),
                ("caf_load", mcf.load_caf_from_sroh_structure)]
    for tag, fn in builders:
        m = fn(verbose=False)
        for k in ("energies", "tdm", "zeeman_x", "zeeman_y", "zeeman_z"):
            out[f"{tag}_{k}"] = np.asarray(getattr(m, k))
        out[f"{tag}_scalars"] = np.array([m.Gamma, m.k, m.mass, m.wavelength])
    # the builders' intermediate matrices, per molecule (same steps as the builders)
    consts = {"sroh": (msr.BX * msr.CM_TO_HZ, msr.DX * msr.CM_TO_HZ, msr.gammaX * msr.CM_TO_HZ, msr.bF_X,
                       msr.Aso * msr.CM_TO_HZ, msr.Be_A * msr.CM_TO_HZ,
                       (msr.p_A + 2 * msr.q_A) * msr.CM_TO_HZ, msr.T_A * msr.CM_TO_HZ),
              "caoh": (mca.BX_HZ, mca.DX_HZ, mca.gammaX_HZ, mca.bF_X_HZ, mca.Aso_A * mca.CM_TO_HZ,
                       mca.Be_A * mca.CM_TO_HZ, (mca.p_A + 2 * mca.q_A) * mca.CM_TO_HZ,
                       mca.T_A * mca.CM_TO_HZ),
              "caf": (mcf.BX_HZ, mcf.DX_HZ, mcf.gammaX_HZ, mcf.bF_X_HZ, mcf.Aso_A * mcf.CM_TO_HZ,
                      mcf.Be_A * mcf.CM_TO_HZ, (mcf.p_A + 2 * mcf.q_A) * mcf.CM_TO_HZ,
                      mcf.T_A * mcf.CM_TO_HZ)}
    gb = enumerate_states(cb.HundsCaseB_LinearMolecule, {"v1": 0, "v2": 0, "v3": 0, "S": 0.5, "I": 0.5,
                                                         "Lambda": 0, "ell": 0, "N": [1]})
    ea = [s for s in enumerate_states(ca.HundsCaseA_LinearMolecule,
                                      {"v1": 0, "v2": 0, "v3": 0, "ell": 0, "Lambda": [-1, 1],
                                       "I": 0.5, "S": 0.5, "J": [0.5]})
          if abs(Fraction(s.P).limit_denominator(10)) == Fraction(1, 2)]
    eb = [s for s in enumerate_states(cb.HundsCaseB_LinearMolecule,
                                      {"v1": 0, "v2": 0, "v3": 0, "S": 0.5, "I": 0.5,
                                       "Lambda": [-1, 0, 1], "ell": 0, "N": [0, 1]})
          if abs(s.J - 0.5) < 0.01 and s.N >= abs(s.K)]
    qn = lambda basis, names: np.array([[getattr(s, n) for n in names] for s in basis])  # noqa: E731
    out["ground_basis_qn"] = qn(gb, ("N", "J", "F", "M"))
    out["excited_a_qn"] = qn(ea, ("Lambda", "Sigma", "J", "P", "F", "M"))
    out["excited_b_qn"] = qn(eb, ("Lambda", "K", "N", "J", "F", "M"))
    for tag, (B, D, g, bF, Aso, Be, p2q, TA) in consts.items():
        H = Hamiltonian(gb)
        H.add_operator("B", B, cb.Rotation)
        H.add_operator("D", D, cb.RotationDistortion)
        H.add_operator("gamma", g, cb.SpinRotation)
        H.add_operator("bF", bF, cb.Hyperfine_IS)
        H.evaluate()
        st = H.solve()
        out[f"{tag}_Hg"] = H.matrix.copy()
        out[f"{tag}_Cg"] = np.array([x.coeffs for x in st])
        He = Hamiltonian(ea)
        He.add_operator("Aso", Aso, ca.SpinOrbit)
        He.add_operator("Be", Be, ca.Rotation)
        He.add_operator("p+2q", p2q, ca.LambdaDoubling_p2q)
        He.evaluate()
        for i in range(len(ea)):
            He.matrix[i, i] += TA
        est = sorted(He.solve(), key=lambda x: x.E)
        out[f"{tag}_He"] = He.matrix.copy()
        out[f"{tag}_Ee_all"] = np.array([x.E for x in est])
        out[f"{tag}_Ce"] = np.array([x.coeffs for x in est[4:]])
    for p in (-1, 0, 1):
        out[f"ground_zeeman_p{p + 1}"] = operator_to_matrix(gb, cb.Zeeman, p)
        out[f"ground_zeeman_nuclear_p{p + 1}"] = operator_to_matrix(gb, cb.zeeman_nuclear, p)
        out[f"tdm_basis_p{p + 1}"] = np.array([[cb.TDM(a, b, p) for b in eb] for a in gb])
    out["overlap_ba"] = np.array([[overlap_caseb_casea(b, a) for a in ea] for b in eb])
    # operators on larger bases
    big_b = enumerate_states(cb.HundsCaseB_LinearMolecule, {"v1": 0, "v2": 0, "v3": 0, "S": 0.5, "I": 0.5,
                                                            "Lambda": 0, "ell": 0, "N": [0, 1, 2]})
    out["big_b_qn"] = qn(big_b, ("N", "J", "F", "M"))
    for name in ("Rotation", "RotationDistortion", "SpinRotation", "Hyperfine_IS"):
        out[f"big_b_{name}"] = operator_to_matrix(big_b, getattr(cb, name))
    for p in (-1, 0, 1):
        out[f"big_b_Zeeman_p{p + 1}"] = operator_to_matrix(big_b, cb.Zeeman, p)
        out[f"big_b_zeeman_nuclear_p{p + 1}"] = operator_to_matrix(big_b, cb.zeeman_nuclear, p)
    big_e = [s for s in enumerate_states(cb.HundsCaseB_LinearMolecule,
                                         {"v1": 0, "v2": 0, "v3": 0, "S": 0.5, "I": 0.5,
                                          "Lambda": [-1, 0, 1], "ell": 0, "N": [0, 1, 2]})
             if s.N >= abs(s.K)]
    out["big_e_qn"] = qn(big_e, ("Lambda", "K", "N", "J", "F", "M"))
    for p in (-1, 0, 1):
        out[f"big_tdm_p{p + 1}"] = np.array([[cb.TDM(a, b, p) for b in big_e] for a in big_b])
    big_a = [s for s in enumerate_states(ca.HundsCaseA_LinearMolecule,
                                         {"v1": 0, "v2": 0, "v3": 0, "ell": 0, "Lambda": [-1, 1],
                                          "I": 0.5, "S": 0.5, "J": [0.5, 1.5]})
             if abs(s.P) <= s.J]
    out["big_a_qn"] = qn(big_a, ("Lambda", "Sigma", "J", "P", "F", "M"))
    for name in ("Rotation", "SpinOrbit", "LambdaDoubling_p2q"):
        out[f"big_a_{name}"] = np.array([[getattr(ca, name)(a, b) for b in big_a] for a in big_a])
    out["big_overlap_ba"] = np.array([[overlap_caseb_casea(b, a) for a in big_a] for b in big_e])
    print("tables:", {k: v.shape for k, v in out.items() if k.startswith("big")})
    np.savez_compressed(os.path.join(HERE, "ref_tables.npz"), **out)


def main():
    tables()
    MolecularData, load_sroh, osd, s3d, DCMOTSimulator, trj, sto = _import_reference()
    sroh = load_sroh(os.path.join(REF, "julia_sim"), verbose=False)
    subdoppler(MolecularData, sroh, osd)
    force3d(MolecularData, sroh, osd, s3d)
    force3d_axis(sroh, osd, s3d)
    traj3d_mc(sroh, DCMOTSimulator, trj)
    sse(sroh, osd, sto)
    analysis()
    floquet(MolecularData, sroh, osd)


if __name__ == "__main__":
    main()
