"""FP64 numpy restatement of molmot's standing-wave Lindblad force.
TEST INFRASTRUCTURE ONLY (the checker of the gfx950 Lindblad kernels).

* ``hamiltonian``       molmot/obe/obe_subdoppler.py:19-104 (build_H_standing_wave),
                        with an optional complex polarisation vector per beam
* ``liouvillian``       :107-141 (build_liouvillian_fast), column-major vec
* ``steady_state``      :144-183 (lstsq on the trace-augmented system,
                        Hermitise, clip negative eigenvalues, normalise)
* ``force_from_rho``    :186-228
* ``force_subdoppler``  :231-286 (average over n_spatial standing-wave phases)
* ``rk4_force``         the time-domain propagator of the device kernels:
                        RK4 on d rho/dt = -i[H(t), rho] + D(rho) with the
                        standing-wave phase moving as phi0 + k v t, force
                        averaged over the second half of the run.

Pinned by: the two-level optical Bloch steady state (closed form), trace and
Hermiticity invariants, and agreement of rk4_force's long-time limit with
steady_state at v = 0 (tests/test_oracle_lindblad.py).
"""
import numpy as np

HBAR = 1.054571817e-34


def hamiltonian(mol, beam_pairs, v, z_macro, z_micro, B_gradient, Omega_Gamma, pol=None):
    """H (rad/s) in the multi-rotating frame; ``pol`` (n_pairs, 2, 3) complex
    replaces the integer polarisations by vectors over q (then the coupling
    uses the complex dipole sum_q eps_q d_q instead of |d_{q}|)."""
    G = mol.Gamma
    ng, ne = mol.n_ground, mol.n_excited
    N = ng + ne
    B_gauss = B_gradient * z_macro * 1e4
    Es = mol.energies + B_gauss * mol.zeeman_z_diag * G / (2 * np.pi)
    Ee_mean = np.mean(Es[ng:])
    lf = [bp["freq"] for bp in beam_pairs]
    Om0 = Omega_Gamma * G
    d = mol.tdm
    H = np.zeros((N, N), dtype=complex)
    for ig in range(ng):
        om = Ee_mean - Es[ig]
        deltas = [f - om for f in lf]
        H[ig, ig] = -2 * np.pi * min(deltas, key=abs)
    for ie in range(ne):
        H[ng + ie, ng + ie] = 2 * np.pi * (Es[ng + ie] - Ee_mean)
    for p, bp in enumerate(beam_pairs):
        Om = Om0 * np.sqrt(bp["s0"])
        for ig in range(ng):
            for ie in range(ne):
                ea = ng + ie
                for side, sign in ((0, 1.0), (1, -1.0)):
                    if pol is None:
                        q = bp["q_fwd"] if side == 0 else bp["q_bwd"]
                        dv = abs(d[ig, ea, q])
                        if dv <= 1e-10:
                            continue
                    else:
                        dv = np.sum(pol[p, side] * d[ig, ea, :])
                        if abs(dv) <= 1e-10:
                            continue
                    c = Om * dv / 2 * np.exp(sign * 1j * z_micro)
                    H[ea, ig] += c
                    H[ig, ea] += np.conj(c)
    return H


def decay_tables(mol, Gamma_eff):
    """Per-excited decay rate and ground repopulation rates of the
    reference's jump operators sqrt(Gamma br_{g,e,q}) |g><e| with br >= 1e-15
    (obe_subdoppler.py:124-139)."""
    ng, ne = mol.n_ground, mol.n_excited
    d2 = mol.d_squared
    rep = np.zeros((ng, ne))
    for ie in range(ne):
        total = sum(d2[ig, ie, q] for ig in range(ng) for q in range(3))
        if total < 1e-30:
            continue
        for ig in range(ng):
            for q in range(3):
                br = d2[ig, ie, q] / total
                if br < 1e-15:
                    continue
                rep[ig, ie] += Gamma_eff * br
    gam = rep.sum(0)
    return rep, gam


def liouvillian(H, mol, Gamma_eff):
    N = H.shape[0]
    ng, ne = mol.n_ground, mol.n_excited
    I = np.eye(N, dtype=complex)
    L = -1j * (np.kron(I, H) - np.kron(H.T, I))
    d2 = mol.d_squared
    for ie in range(ne):
        ea = ng + ie
        total = sum(d2[ig, ie, q] for ig in range(ng) for q in range(3))
        if total < 1e-30:
            continue
        for ig in range(ng):
            for q in range(3):
                br = d2[ig, ie, q] / total
                if br < 1e-15:
                    continue
                Lc = np.zeros((N, N), dtype=complex)
                Lc[ig, ea] = np.sqrt(Gamma_eff * br)
                LdL = Lc.conj().T @ Lc
                L += np.kron(Lc.conj(), Lc)
                L -= 0.5 * np.kron(I, LdL)
                L -= 0.5 * np.kron(LdL.T, I)
    return L


def steady_state(L, N):
    trace_row = np.zeros((1, N * N), dtype=complex)
    for i in range(N):
        trace_row[0, i * N + i] = 1.0
    A = np.vstack([L, trace_row])
    rhs = np.zeros(N * N + 1, dtype=complex)
    rhs[-1] = 1.0
    x = np.linalg.lstsq(A, rhs, rcond=None)[0]
    rho = x.reshape((N, N), order="F")
    rho = 0.5 * (rho + rho.conj().T)
    w, U = np.linalg.eigh(rho)
    rho = (U * np.maximum(w, 0)) @ U.conj().T
    tr = np.trace(rho)
    return rho / tr if abs(tr) > 1e-15 else np.eye(N, dtype=complex) / N


def force_from_rho(rho, mol, H_fwd_coupling, H_bwd_coupling):
    """F = sum 2 hbar k Im(rho_ge c_eg) (fwd) - (bwd)  (obe_subdoppler.py:186-228);
    the couplings are the (e, g) blocks of each beam direction."""
    ng = mol.n_ground
    blk = rho[:ng, ng:]                      # rho[g, e]
    f = 2 * HBAR * mol.k * (np.imag(np.sum(blk * H_fwd_coupling[ng:, :ng].T))
                            - np.imag(np.sum(blk * H_bwd_coupling[ng:, :ng].T)))
    return f


def coupling_blocks(mol, beam_pairs, Omega_Gamma, pol=None):
    """(A, B): fwd / bwd coupling matrices (rad/s) without the phase factor,
    nonzero in the (e, g) block only."""
    zero = dict(freq=0.0)
    mol_B0 = mol
    ng, ne = mol.n_ground, mol.n_excited
    N = ng + ne
    Om0 = Omega_Gamma * mol.Gamma
    d = mol.tdm
    A = np.zeros((N, N), dtype=complex)
    B = np.zeros((N, N), dtype=complex)
    for p, bp in enumerate(beam_pairs):
        Om = Om0 * np.sqrt(bp["s0"])
        for ig in range(ng):
            for ie in range(ne):
                ea = ng + ie
                for side, M in ((0, A), (1, B)):
                    if pol is None:
                        q = bp["q_fwd"] if side == 0 else bp["q_bwd"]
                        dv = abs(d[ig, ea, q])
                        if dv <= 1e-10:
                            continue
                    else:
                        dv = np.sum(pol[p, side] * d[ig, ea, :])
                        if abs(dv) <= 1e-10:
                            continue
                    M[ea, ig] += Om * dv / 2
    del zero, mol_B0
    return A, B


def force_subdoppler(mol, beam_pairs, v, z_macro, B_gradient, Omega_Gamma, n_spatial=16,
                     Gamma_eff_factor=1.0, pol=None):
    """(F, rho_avg, R) averaged over n_spatial phases (obe_subdoppler.py:231-286)."""
    N = mol.n_states
    Geff = mol.Gamma * Gamma_eff_factor
    A, B = coupling_blocks(mol, beam_pairs, Omega_Gamma, pol)
    F = 0.0
    rho_avg = np.zeros((N, N), dtype=complex)
    for zm in np.linspace(0, 2 * np.pi, n_spatial, endpoint=False):
        H = hamiltonian(mol, beam_pairs, v, z_macro, zm, B_gradient, Omega_Gamma, pol)
        rho = steady_state(liouvillian(H, mol, Geff), N)
        F += force_from_rho(rho, mol, A * np.exp(1j * zm), B * np.exp(-1j * zm))
        rho_avg += rho
    F /= n_spatial
    rho_avg /= n_spatial
    R = Geff * np.sum(np.real(np.diag(rho_avg))[mol.n_ground:])
    return F, rho_avg, R


def rk4_force(mol, beam_pairs, v, phi0, z_macro, B_gradient, Omega_Gamma, t_total, dt,
              Gamma_eff_factor=1.0, pol=None, avg_from=0.5, rho0=None):
    """RK4 propagation in natural units (t Gamma, H / Gamma) from rho0 (default:
    ground states equally populated).  The standing-wave phase is
    phi(t) = phi0 + (k v / Gamma) t.  Returns (F_avg [N], R_avg [1/s], rho_T):
    F and the excited population averaged over the samples at the end of
    every step with t > avg_from * t_total."""
    G = mol.Gamma
    ng = mol.n_ground
    N = mol.n_states
    Hd = np.real(np.diag(hamiltonian(mol, beam_pairs, 0.0, z_macro, 0.0, B_gradient, 0.0))) / G
    A, B = coupling_blocks(mol, beam_pairs, Omega_Gamma, pol)
    A, B = A / G, B / G
    rep, gam = decay_tables(mol, Gamma_eff_factor)
    gvec = np.concatenate([np.zeros(ng), gam])
    kv = mol.k * v / G
    if rho0 is None:
        rho0 = np.zeros((N, N), dtype=complex)
        rho0[np.arange(ng), np.arange(ng)] = 1.0 / ng

    def H_at(t):
        ph = phi0 + kv * t
        C = np.exp(1j * ph) * A + np.exp(-1j * ph) * B
        return np.diag(Hd).astype(complex) + C + C.conj().T, ph

    def rhs(t, r):
        H, _ = H_at(t)
        X = H @ r
        d = -1j * (X - X.conj().T)
        d -= 0.5 * (gvec[:, None] + gvec[None, :]) * r
        pe = np.real(np.diag(r))[ng:]
        d[np.arange(ng), np.arange(ng)] += rep @ pe
        return d

    n_steps = int(round(t_total / dt))
    r = rho0.copy()
    Fs, Rs, cnt = 0.0, 0.0, 0
    for s in range(n_steps):
        t = s * dt
        k1 = rhs(t, r)
        k2 = rhs(t + dt / 2, r + dt / 2 * k1)
        k3 = rhs(t + dt / 2, r + dt / 2 * k2)
        k4 = rhs(t + dt, r + dt * k3)
        r = r + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        if (s + 1) * dt > avg_from * t_total:
            ph = phi0 + kv * (s + 1) * dt
            blk = r[:ng, ng:]
            f = 2 * HBAR * mol.k * G * (np.imag(np.sum(blk * (np.exp(1j * ph) * A[ng:, :ng]).T))
                                         - np.imag(np.sum(blk * (np.exp(-1j * ph) * B[ng:, :ng]).T)))
            Fs += f
            Rs += G * Gamma_eff_factor * np.sum(np.real(np.diag(r))[ng:])
            cnt += 1
    return Fs / max(cnt, 1), Rs / max(cnt, 1), r
