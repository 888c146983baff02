"""FP64 numpy restatement of molmot's Floquet OBE force.
TEST INFRASTRUCTURE ONLY (the checker of the gfx950 Floquet solver).

* ``blocks``        FloquetOBE._build_floquet_blocks (molmot/obe/floquet.py:
                    78-192): H0 (nearest-laser ground detunings, excited
                    energies about their mean), H_plus / H_minus, and the
                    Liouvillian blocks L0 (with the dissipator), L_plus,
                    L_minus, column-major vec
* ``solve``         _solve_floquet (:194-300): the dense block-tridiagonal
                    system with (L0 - i n kv) on the diagonal, L_plus below
                    and L_minus above, the last row of the n = 0 block
                    replaced by the trace, and the Zeeman diagonal at z_macro
* ``force``         force_vs_velocity (:302-378): F = hbar k (Im Tr(rho_-1
                    H_plus) - Im Tr(rho_+1 H_minus)), v = 0 replaced by 1e-6;
                    for the standing-wave model 2 hbar k (Im Tr(rho_-1 A) -
                    Im Tr(rho_+1 B)), A / B the forward / backward (e, g)
                    couplings
* ``scattering_rate``  (:380-403)

``coupling="reference"`` is the reference's model, in which H_plus and
H_minus are each Hermitian (both (e, g) and (g, e) entries carry the
forward, resp. backward, coupling).  ``coupling="standing_wave"`` is the
Hamiltonian of obe_subdoppler.build_H_standing_wave (:19-104), H(z) = H0 +
C e^{ikz} + C^dagger e^{-ikz} with C = A + B^dagger, whose time-periodic
steady state the RK4 propagator reaches; its force is the time average of
2 hbar k Im sum rho_ge (e^{ikz} A_eg - e^{-ikz} B_eg).

Pinned by (tests/test_oracle_floquet.py): the reference's own outputs
(tests/golden/ref_floquet.npz); for the standing-wave model, the time
average of oracle.lindblad.rk4_force over whole periods of a moving atom
(2e-6) and, loosely, the adiabatic limit v -> 0 (the phase average of the
steady-state force).
"""
import numpy as np

HBAR = 1.054571817e-34


def _dissipator(mol, Gamma_eff, N):
    ng, ne = mol.n_ground, mol.n_excited
    I = np.eye(N, dtype=complex)
    L = np.zeros((N * N, N * N), dtype=complex)
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
                L += np.kron(Lc.conj(), Lc) - 0.5 * np.kron(I, LdL) - 0.5 * np.kron(LdL.T, I)
    return L


def blocks(mol, beam_pairs, Omega_Gamma=1.0, Gamma_eff_factor=1.0, coupling="reference"):
    """dict(H0, H_plus, H_minus, L0, L_plus, L_minus) in rad/s."""
    ng, ne = mol.n_ground, mol.n_excited
    N = ng + ne
    G = mol.Gamma
    E = mol.energies.copy()
    Ee = np.mean(E[ng:])
    lf = [bp["freq"] for bp in beam_pairs]
    H0 = np.zeros((N, N), dtype=complex)
    for ig in range(ng):
        om = Ee - E[ig]
        H0[ig, ig] = -2 * np.pi * min([f - om for f in lf], key=abs)
    for ie in range(ne):
        H0[ng + ie, ng + ie] = 2 * np.pi * (E[ng + ie] - Ee)
    Hp = np.zeros((N, N), dtype=complex)
    Hm = np.zeros((N, N), dtype=complex)
    A = np.zeros((N, N), dtype=complex)   # forward coupling, (e, g) block only
    B = np.zeros((N, N), dtype=complex)   # backward coupling, (e, g) block only
    Om0 = Omega_Gamma * G
    d = mol.tdm
    for bp in beam_pairs:
        Om = Om0 * np.sqrt(bp["s0"])
        for ig in range(ng):
            for ie in range(ne):
                ea = ng + ie
                df = abs(d[ig, ea, bp["q_fwd"]])
                if df > 1e-10:
                    c = Om * df / 2
                    A[ea, ig] += c
                    Hp[ea, ig] += c
                    if coupling == "reference":
                        Hp[ig, ea] += np.conj(c)
                    else:
                        Hm[ig, ea] += np.conj(c)
                db = abs(d[ig, ea, bp["q_bwd"]])
                if db > 1e-10:
                    c = Om * db / 2
                    B[ea, ig] += c
                    Hm[ea, ig] += c
                    if coupling == "reference":
                        Hm[ig, ea] += np.conj(c)
                    else:
                        Hp[ig, ea] += np.conj(c)
    I = np.eye(N, dtype=complex)
    L0 = -1j * (np.kron(I, H0) - np.kron(H0.T, I)) + _dissipator(mol, G * Gamma_eff_factor, N)
    Lp = -1j * (np.kron(I, Hp) - np.kron(Hp.T, I))
    Lm = -1j * (np.kron(I, Hm) - np.kron(Hm.T, I))
    return dict(H0=H0, H_plus=Hp, H_minus=Hm, L0=L0, L_plus=Lp, L_minus=Lm, N=N, A=A, B=B, coupling=coupling)


def solve(mol, blk, v, n_max, B_gradient=0.0, z_macro=0.0):
    """Fourier components rho_{-n_max} .. rho_{n_max}, each (N, N)."""
    N = blk["N"]
    N2 = N * N
    nb = 2 * n_max + 1
    kv = mol.k * v
    L0 = blk["L0"]
    if abs(z_macro) > 1e-15:
        Bg = B_gradient * z_macro * 1e4
        EZ = Bg * mol.zeeman_z_diag * mol.Gamma / (2 * np.pi)
        I = np.eye(N, dtype=complex)
        L0 = L0 + -1j * (np.kron(I, np.diag(2 * np.pi * EZ)) - np.kron(np.diag(2 * np.pi * EZ), I))
    M = np.zeros((nb * N2, nb * N2), dtype=complex)
    for b in range(nb):
        n = b - n_max
        r = slice(b * N2, (b + 1) * N2)
        M[r, r] = L0 - 1j * n * kv * np.eye(N2)
        if b > 0:
            M[r, (b - 1) * N2:b * N2] = blk["L_plus"]
        if b < nb - 1:
            M[r, (b + 1) * N2:(b + 2) * N2] = blk["L_minus"]
    row = n_max * N2 + N2 - 1
    M[row, :] = 0.0
    for i in range(N):
        M[row, n_max * N2 + i * N + i] = 1.0
    rhs = np.zeros(nb * N2, dtype=complex)
    rhs[row] = 1.0
    try:
        x = np.linalg.solve(M, rhs)
    except np.linalg.LinAlgError:
        x = np.linalg.lstsq(M, rhs, rcond=None)[0]
    return [x[b * N2:(b + 1) * N2].reshape((N, N), order="F") for b in range(nb)]


def _veff(v):
    return v if abs(v) >= 1e-12 else 1e-6


def force(mol, blk, v_arr, n_max, B_gradient=0.0, z_macro=0.0, hbar=HBAR):
    F = np.zeros(len(v_arr))
    for i, v in enumerate(v_arr):
        rc = solve(mol, blk, _veff(v), n_max, B_gradient, z_macro)
        rm1 = rc[n_max - 1] if n_max >= 1 else np.zeros_like(rc[0])
        rp1 = rc[n_max + 1] if n_max >= 1 else np.zeros_like(rc[0])
        if blk["coupling"] == "reference":
            F[i] = hbar * mol.k * (np.imag(np.trace(rm1 @ blk["H_plus"])) - np.imag(np.trace(rp1 @ blk["H_minus"])))
        else:
            # time average of 2 hbar k Im sum rho_ge (e^{ikvt} A_eg - e^{-ikvt} B_eg)
            F[i] = 2 * hbar * mol.k * (np.imag(np.trace(rm1 @ blk["A"])) - np.imag(np.trace(rp1 @ blk["B"])))
    return F


def scattering_rate(mol, blk, v, n_max, Gamma_eff_factor=1.0, B_gradient=0.0, z_macro=0.0):
    rho0 = solve(mol, blk, _veff(v), n_max, B_gradient, z_macro)[n_max]
    return max(mol.Gamma * Gamma_eff_factor * np.sum(np.real(np.diag(rho0))[mol.n_ground:]), 0.0)
