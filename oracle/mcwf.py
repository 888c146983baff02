"""FP64 numpy restatement of molmot's Monte Carlo wavefunction (quantum jump)
solver.  TEST INFRASTRUCTURE ONLY: the checker of the gfx950 MCWF kernel.

* ``wigner_d1`` / ``rotate_polarization``  molmot/obe/stochastic.py:59-156
* ``tables``          SSEProblem.__init__ / _setup_beam_params (:321-440)
* ``omega0s_zeeman``  get_omega0s_zeeman (:476-524), ``b_field`` (:526-549)
* ``fields``          compute_fields_6beam (:173-277)
* ``rhs``             SSESolver._rhs (:806-954), literally: interaction
                      picture with the absolute phases exp(-i omega0 t) and
                      the laser phases omega_f t
* ``rk4_step``        SSESolver._rk4_step (:956-981)
* ``run``             the step loop of SSESolver.run (:661-804) with the
                      jump logic of the device reference kernel
                      (molmot_cuda/stochastic_cuda.py:600-748: cumulative
                      channel search, optional escape and renormalisation)

``shift`` = (w_ref_g, w_ref_e) measures ground and excited energies from
common references and the laser frequencies from w_ref_e - w_ref_g; every
phase that enters the equations is a difference omega_f - omega_e + omega_g or
omega_i - omega_j within one manifold, so the shift changes nothing but the
rounding of the huge absolute phases (tests pin the equivalence).

Random numbers: Philox4x32-10 with key (seed_lo, seed_hi) and counter
(block, stream, trajectory, 0); block 0 = first threshold; jump m (1-based)
uses blocks 3m-2 (channel, new threshold), 3m-1 (recoil direction) and 3m
(diffusion signs); renormalisation redraws use stream 1.  Two 32-bit words
(a, b) give u = ((a >> 5) 2^26 + (b >> 6) + 1/2) 2^-53 in (0, 1).

Pinned by (tests/test_oracle_mcwf.py): two-level no-jump evolution against
the matrix exponential of the non-Hermitian Hamiltonian; field-free Zeeman
precession against expm; the force against -d<H>/dr by finite differences;
d|psi|^2/dt = -sum_e |psi_e|^2; rotation unitarity and the helicity flip of
the backward z beam; shifted vs literal phases.
"""
import numpy as np

from .philox import philox4x32_10

HBAR = 1.054571817e-34
_DIRS = (np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, 1.0]),
         -np.array([1.0, 0.0, 0.0]), -np.array([0.0, 1.0, 0.0]), -np.array([0.0, 0.0, 1.0]))


def wigner_d1(cb, sb, alpha, gamma):
    """j = 1 Wigner D (ZYZ), rows/cols q = -1, 0, +1 (stochastic.py:59-88)."""
    em_a, ep_a = np.exp(-1j * alpha), np.exp(1j * alpha)
    em_g, ep_g = np.exp(-1j * gamma), np.exp(1j * gamma)
    r2 = np.sqrt(2)
    return np.array([[0.5 * (1 + cb) * em_a * em_g, -sb / r2 * em_a, 0.5 * (1 - cb) * em_a * ep_g],
                     [sb / r2 * em_g, cb, -sb / r2 * ep_g],
                     [0.5 * (1 - cb) * ep_a * em_g, sb / r2 * ep_a, 0.5 * (1 + cb) * ep_a * ep_g]],
                    dtype=complex)


def euler_angles(k_hat):
    """(alpha, beta, gamma) of stochastic.py:125-150, or None for k_hat = +z:
    the ZYZ angles read off the Rodrigues matrix about the axis k_hat x z
    with its y component negated and its length sin(theta) kept."""
    k_hat = np.asarray(k_hat, float)
    k_hat = k_hat / np.linalg.norm(k_hat)
    ct = np.clip(k_hat[2], -1.0, 1.0)
    th = np.arccos(ct)
    if abs(th) < 1e-12:
        return None
    if abs(th - np.pi) < 1e-12:
        return 0.0, np.pi, 0.0
    x, y, z = np.cross(k_hat, [0.0, 0.0, 1.0])
    y = -y
    st, c = np.sin(th), np.cos(th)
    A33 = (1 - c) * z * z + c
    A31 = (1 - c) * z * x - y * st
    A32 = (1 - c) * z * y + x * st
    A13 = (1 - c) * x * z + y * st
    A23 = (1 - c) * y * z - x * st
    return np.arctan2(A23, A13), np.arctan2(np.sqrt(max(0.0, 1 - A33 * A33)), A33), np.arctan2(A32, -A31)


def rotate_polarization(pol, k_hat):
    """stochastic.py:103-156: Euler angles of the rotation taking z to k_hat
    (axis k_hat x z with its y component negated), then solve D x = pol."""
    ang = euler_angles(k_hat)
    if ang is None:
        return np.array(pol, dtype=complex)
    a, b, g = ang
    return np.linalg.solve(wigner_d1(np.cos(b), np.sin(b), a, g), np.asarray(pol, dtype=complex))


def tables(mol, beam_pairs, B_gradient, beam_radius=10e-3, add_spontaneous_kick=True,
           diffusion_constants=None, hbar=HBAR):
    """Natural-unit problem tables of SSEProblem (stochastic.py:321-440)."""
    G, k, m = mol.Gamma, mol.k, mol.mass
    ng, ne = mol.n_ground, mol.n_excited
    d = mol.tdm
    d_ge = np.zeros((ng, ne, 3))
    for ig in range(ng):
        for ie in range(ne):
            for q in range(3):
                d_ge[ig, ie, q] = abs(d[ig, ng + ie, q])
    br = np.zeros((ne, ng, 3))
    for ie in range(ne):
        tot = 0.0
        for ig in range(ng):
            for q in range(3):
                tot += d_ge[ig, ie, q] ** 2
        if tot > 1e-30:
            for ig in range(ng):
                for q in range(3):
                    br[ie, ig, q] = d_ge[ig, ie, q] ** 2 / tot
    nf = len(beam_pairs)
    wl = np.array([2.0 * np.pi * bp["freq"] / G for bp in beam_pairs])
    sats = np.array([float(bp["s0"]) for bp in beam_pairs])
    pols = []
    for bp in beam_pairs:
        p = np.zeros(3, dtype=complex)
        p[bp["q_fwd"]] = 1.0
        pols.append(p)
    eps = np.zeros((6, nf, 3), dtype=complex)
    for i, p in enumerate(pols):
        for b, kd in enumerate(_DIRS):
            eps[b, i] = rotate_polarization(p, kd)
    v_nat = G / k
    dif = np.zeros(3) if diffusion_constants is None else \
        np.asarray(diffusion_constants, float) / ((hbar * k) ** 2 * G)
    return dict(ng=ng, ne=ne, N=ng + ne, omega0s=2.0 * np.pi * mol.energies / G, d_ge=d_ge,
                branching=br, omega_lasers=wl, sats=sats, eps=eps,
                zx=np.array(mol.zeeman_x[:ng, :ng], dtype=complex),
                zy=np.array(mol.zeeman_y[:ng, :ng], dtype=complex),
                zz=np.array(mol.zeeman_z[:ng, :ng], dtype=complex),
                zdiag=np.asarray(mol.zeeman_z_diag, float), B_gradient=float(B_gradient),
                r_nat=1.0 / k, v_nat=v_nat, t_nat=1.0 / G, F_nat=hbar * k * G,
                mass_nat=m * G / (hbar * k * k), v_recoil_nat=(hbar * k / m) / v_nat,
                beam_radius_k=beam_radius * k, add_kick=bool(add_spontaneous_kick), diffusion_nat=dif)


def default_shift(tab):
    """(mean ground, mean excited) natural angular frequencies."""
    w = tab["omega0s"]
    return float(np.mean(w[:tab["ng"]])), float(np.mean(w[tab["ng"]:]))


def b_field(tab, r):
    rs = r * tab["r_nat"]
    Bp = tab["B_gradient"]
    return np.array([-Bp * rs[0] / 2.0 * 1e4, -Bp * rs[1] / 2.0 * 1e4, Bp * rs[2] * 1e4])


def omega0s_zeeman(tab, r, shift=(0.0, 0.0)):
    ng = tab["ng"]
    Bx, By, Bz = b_field(tab, r)
    w = tab["omega0s"].copy()
    w[:ng] -= shift[0]
    w[ng:] -= shift[1]
    z = np.zeros(tab["N"])
    z[:ng] = np.real(Bx * np.diag(tab["zx"]) + By * np.diag(tab["zy"]) + Bz * np.diag(tab["zz"]))
    z[ng:] = Bz * tab["zdiag"][ng:]
    return w + z


def fields(tab, r, t, shift=(0.0, 0.0)):
    """E_kq (6, 3) (stochastic.py:173-277)."""
    wl = tab["omega_lasers"] - (shift[1] - shift[0])
    denom = tab["beam_radius_k"] ** 2 / 2.0
    rp = np.array([r[1] ** 2 + r[2] ** 2, r[0] ** 2 + r[2] ** 2, r[0] ** 2 + r[1] ** 2])
    gauss = np.sqrt(np.exp(-rp / denom))
    nf = len(wl)
    a = np.zeros((6, nf), dtype=complex)
    for f in range(nf):
        Gf = np.sqrt(tab["sats"][f]) / (2.0 * np.sqrt(2.0))
        wt = wl[f] * t
        for ax in range(3):
            a[ax, f] = Gf * gauss[ax] * np.exp(1j * (-r[ax] + wt))
            a[ax + 3, f] = Gf * gauss[ax] * np.exp(1j * (r[ax] + wt))
    E = np.zeros((6, 3), dtype=complex)
    for q in range(3):
        for b in range(6):
            for f in range(nf):
                E[b, q] += a[b, f] * np.conj(tab["eps"][b, f, q])
    return E


def rhs(tab, t, psi, r, v, shift=(0.0, 0.0)):
    """(dpsi, F) for unnormalised psi (stochastic.py:806-954)."""
    ng = tab["ng"]
    w = omega0s_zeeman(tab, r, shift)
    pint = psi * np.exp(-1j * w * t)
    pg, pe = pint[:ng], pint[ng:]
    d = tab["d_ge"]
    psi_q_g = np.einsum("geq,e->gq", d, pe)
    psi_q_e = np.einsum("geq,g->eq", d, pg)
    E = fields(tab, r, t, shift)
    Et = E.sum(0)
    d_exp = np.conj(pg) @ psi_q_g
    F = np.zeros(3)
    for k in range(3):
        Fk = 0.0
        for q in range(3):
            dre = E[k, q].real - E[k + 3, q].real
            dim = E[k, q].imag - E[k + 3, q].imag
            Fk -= 2.0 * (d_exp[q].real * dim - d_exp[q].imag * (-dre))
        F[k] = Fk
    dpi = np.zeros(tab["N"], dtype=complex)
    dpi[:ng] = -1j * (psi_q_g @ Et)
    dpi[ng:] = -1j * (psi_q_e @ np.conj(Et)) - 0.5 * pe
    Bx, By, Bz = b_field(tab, r)
    HZ = Bx * tab["zx"] + By * tab["zy"] + Bz * tab["zz"]
    ph = np.exp(-1j * w[:ng] * t)
    HZi = HZ * np.outer(ph, np.conj(ph))
    np.fill_diagonal(HZi, 0.0)
    dpi[:ng] += -1j * (HZi @ pg)
    return dpi * np.exp(1j * w * t), F


def rk4_step(tab, t, psi, r, v, dt, shift=(0.0, 0.0)):
    """One RK4 step of (psi, r, v) (stochastic.py:956-981); returns the new
    state and the stage-1 force."""
    m = tab["mass_nat"]
    k1p, F1 = rhs(tab, t, psi, r, v, shift)
    k1r, k1v = v, F1 / m
    p2, r2, v2 = psi + 0.5 * dt * k1p, r + 0.5 * dt * k1r, v + 0.5 * dt * k1v
    k2p, F2 = rhs(tab, t + 0.5 * dt, p2, r2, v2, shift)
    k2r, k2v = v2, F2 / m
    p3, r3, v3 = psi + 0.5 * dt * k2p, r + 0.5 * dt * k2r, v + 0.5 * dt * k2v
    k3p, F3 = rhs(tab, t + 0.5 * dt, p3, r3, v3, shift)
    k3r, k3v = v3, F3 / m
    p4, r4, v4 = psi + dt * k3p, r + dt * k3r, v + dt * k3v
    k4p, F4 = rhs(tab, t + dt, p4, r4, v4, shift)
    k4r, k4v = v4, F4 / m
    c = dt / 6.0
    return (psi + c * (k1p + 2 * k2p + 2 * k3p + k4p), r + c * (k1r + 2 * k2r + 2 * k3r + k4r),
            v + c * (k1v + 2 * k2v + 2 * k3v + k4v), F1)


def uniform_pair(a, b):
    return (((a >> 5) * 67108864.0) + (b >> 6) + 0.5) * (2.0 ** -53)


def draws(seed, traj, block, stream=0):
    w = philox4x32_10((block, stream, traj & 0xFFFFFFFF, traj >> 32), (seed & 0xFFFFFFFF, seed >> 32))
    return uniform_pair(w[0], w[1]), uniform_pair(w[2], w[3]), w


def threshold(u, rule):
    """rule 0: -ln u (the reference, stochastic.py:742-746); rule 1: 1 - u
    (jump when |psi|^2 < u)."""
    return 1.0 - u if rule else -np.log(u)


def run(tab, r0, v0, psi0, dt, n_steps, save_every, seed, traj=0, shift=(0.0, 0.0), r_escape=0.0,
        renorm_interval=0, avg_start=0, jump_rule=0):
    """One trajectory in natural units.  Returns dict(t, r, v, pop, F,
    n_scatters, escaped, n_saved, jumps=[(step, t, ie, ig, q)], force_sum,
    n_avg, psi, threshold).  F at a save point and force_sum (over steps >=
    avg_start, stage-1 force of each step) use the normalised state."""
    N, ng, ne = tab["N"], tab["ng"], tab["ne"]
    psi = np.array(psi0, dtype=complex)
    r = np.array(r0, float)
    v = np.array(v0, float)
    n_save = n_steps // save_every + 1
    out = dict(t=np.zeros(n_save), r=np.zeros((n_save, 3)), v=np.zeros((n_save, 3)),
               pop=np.zeros((n_save, N)), F=np.zeros((n_save, 3)))
    thr = threshold(draws(seed, traj, 0)[0], jump_rule)
    t = 0.0
    last = 0.0
    jumps = []
    n_sc = 0
    esc = 0
    si = 0
    fsum = np.zeros(3)
    n_avg = 0
    n_jump = 0
    n_ren = 0
    br = tab["branching"]
    for step in range(n_steps + 1):
        if step % save_every == 0 and si < n_save:
            ns = float(np.sum(psi.real ** 2 + psi.imag ** 2))
            inv = 1.0 / ns if ns > 1e-30 else 1.0
            out["t"][si] = t
            out["r"][si] = r
            out["v"][si] = v
            out["pop"][si] = (psi.real ** 2 + psi.imag ** 2) * inv
            out["F"][si] = rhs(tab, t, psi, r, v, shift)[1] * inv
            si += 1
        if step == n_steps:
            break
        if r_escape > 0 and r[0] ** 2 + r[1] ** 2 + r[2] ** 2 > r_escape ** 2:
            esc = 1
            break
        ns0 = float(np.sum(psi.real ** 2 + psi.imag ** 2))
        psi, r, v, F1 = rk4_step(tab, t, psi, r, v, dt, shift)
        if step >= avg_start:
            fsum += F1 / ns0
            n_avg += 1
        t += dt
        ns = float(np.sum(psi.real ** 2 + psi.imag ** 2))
        if 1.0 - ns > thr:
            n_jump += 1
            u_ch, u_thr, _ = draws(seed, traj, 3 * n_jump - 2)
            tot = 0.0
            for ie in range(ne):
                pe = psi[ng + ie].real ** 2 + psi[ng + ie].imag ** 2
                if pe < 1e-30:
                    continue
                for ig in range(ng):
                    for q in range(3):
                        if br[ie, ig, q] < 1e-15:
                            continue
                        tot += br[ie, ig, q] * pe
            if tot > 1e-30:
                rc = u_ch * tot
                cum = 0.0
                sel = None
                for ie in range(ne):
                    pe = psi[ng + ie].real ** 2 + psi[ng + ie].imag ** 2
                    if pe < 1e-30:
                        continue
                    for ig in range(ng):
                        for q in range(3):
                            if br[ie, ig, q] < 1e-15:
                                continue
                            cum += br[ie, ig, q] * pe
                            if sel is None and rc <= cum:
                                sel = (ie, ig, q)
                psi = np.zeros(N, dtype=complex)
                psi[sel[1]] = 1.0
                n_sc += 1
                jumps.append((step, t) + sel)
                if tab["add_kick"]:
                    u_c, u_p, _ = draws(seed, traj, 3 * n_jump - 1)
                    ct = 2.0 * u_c - 1.0
                    st = np.sqrt(1.0 - ct * ct)
                    ph = 2.0 * np.pi * u_p
                    v = v + tab["v_recoil_nat"] * np.array([st * np.cos(ph), st * np.sin(ph), ct])
            else:
                psi = np.zeros(N, dtype=complex)
                psi[0] = 1.0
            dn = t - last
            if np.any(tab["diffusion_nat"] > 0):
                _, _, w = draws(seed, traj, 3 * n_jump)
                for i in range(3):
                    if tab["diffusion_nat"][i] > 0:
                        v[i] += (1.0 if w[i] >= 0x80000000 else -1.0) * np.sqrt(2.0 * tab["diffusion_nat"][i] * dn)
            last = t
            thr = threshold(u_thr, jump_rule)
        if renorm_interval > 0 and (step + 1) % renorm_interval == 0:
            ns = float(np.sum(psi.real ** 2 + psi.imag ** 2))
            if ns > 1.0 + 1e-10 or ns < 1e-15:
                psi = psi * (1.0 / np.sqrt(ns))
                n_ren += 1
                thr = threshold(draws(seed, traj, n_ren, 1)[0], jump_rule)
    out.update(n_scatters=n_sc, escaped=esc, n_saved=si, jumps=jumps, psi=psi, threshold=thr, r_final=r,
               v_final=v, force_sum=fsum, n_avg=n_avg)
    return out


def heff(tab, t, r, shift=(0.0, 0.0)):
    """H_eff (N x N) with dpsi/dt = -i H_eff psi for the linear psi part of
    ``rhs`` at fixed (t, r)."""
    N = tab["N"]
    H = np.zeros((N, N), dtype=complex)
    z = np.zeros(3)
    for j in range(N):
        e = np.zeros(N, dtype=complex)
        e[j] = 1.0
        H[:, j] = 1j * rhs(tab, t, e, r, z, shift)[0]
    return H


def heff_direct(tab, t, r, shift=(0.0, 0.0)):
    """H_eff assembled from its blocks: H[g, e] = sum_q d[g,e,q] Et[q]
    exp(i (theta_g - theta_e)), H[e, g] = conj(H[g, e]), off-diagonal ground
    Zeeman block, H[e, e] = -i/2 (pinned against ``heff``)."""
    ng, N = tab["ng"], tab["N"]
    w = omega0s_zeeman(tab, r, shift)
    th = w * t
    Et = fields(tab, r, t, shift).sum(0)
    H = np.zeros((N, N), dtype=complex)
    cp = np.einsum("geq,q->ge", tab["d_ge"], Et)
    H[:ng, ng:] = cp * np.exp(1j * (th[:ng, None] - th[None, ng:]))
    H[ng:, :ng] = np.conj(H[:ng, ng:]).T
    Bx, By, Bz = b_field(tab, r)
    HZ = Bx * tab["zx"] + By * tab["zy"] + Bz * tab["zz"]
    H[:ng, :ng] = HZ - np.diag(np.diag(HZ))
    H[np.arange(ng, N), np.arange(ng, N)] = -0.5j
    return H


def force_rho(tab, t, r, rho, shift=(0.0, 0.0)):
    """Force of a density matrix: the bilinear form of ``rhs``'s force with
    conj(psi_g) psi_e -> rho[e, g]."""
    ng = tab["ng"]
    w = omega0s_zeeman(tab, r, shift)
    ph = np.exp(-1j * w * t)
    E = fields(tab, r, t, shift)
    d_exp = np.einsum("geq,eg->q", tab["d_ge"], rho[ng:, :ng] * np.outer(ph[ng:], np.conj(ph[:ng])))
    Ed = E[:3] - E[3:]
    return -2.0 * np.imag(np.einsum("q,kq->k", d_exp, Ed))


def lindblad_path(tab, r0, v, rho0, dt, n_steps, shift=(0.0, 0.0)):
    """Density matrix along r(t) = r0 + v t under the master equation whose
    unravelling is the MCWF above: d rho/dt = -i (H rho - rho H^+) +
    sum_{e,g,q} br |g><e| rho |e><g|.  Returns (t (n+1,), F (n+1, 3), pe (n+1,))."""
    ng, ne = tab["ng"], tab["ne"]
    rep = tab["branching"].sum(2)                      # (ne, ng)
    r0 = np.asarray(r0, float)
    v = np.asarray(v, float)

    def f(t, rho):
        H = heff_direct(tab, t, r0 + v * t, shift)
        d = -1j * (H @ rho - rho @ H.conj().T)
        pe = np.real(np.diag(rho))[ng:]
        d[np.arange(ng), np.arange(ng)] += rep.T @ pe
        return d

    rho = np.array(rho0, dtype=complex)
    ts, Fs, pes = [0.0], [force_rho(tab, 0.0, r0, rho, shift)], [float(np.real(np.trace(rho[ng:, ng:])))]
    t = 0.0
    for _ in range(n_steps):
        k1 = f(t, rho)
        k2 = f(t + 0.5 * dt, rho + 0.5 * dt * k1)
        k3 = f(t + 0.5 * dt, rho + 0.5 * dt * k2)
        k4 = f(t + dt, rho + dt * k3)
        rho = rho + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        t += dt
        ts.append(t)
        Fs.append(force_rho(tab, t, r0 + v * t, rho, shift))
        pes.append(float(np.real(np.trace(rho[ng:, ng:]))))
    return np.array(ts), np.array(Fs), np.array(pes)
