"""FP64 numpy restatement of molmot's 3D rate-equation force.
TEST INFRASTRUCTURE ONLY: the checker of the gfx950 3D force kernels.

* ``six_beams``         make_6beam_config (molmot/mot/simulator_3d.py:62-108):
                        per beam pair, beams +x -x +y -y +z -z with q_fwd on
                        the forward and q_bwd on the backward beams
* ``pol_weights``       _decompose_polarization_local (:111-165)
* ``force``             rate_eq_force_3d (:168-365), literally: eigh of the
                        Hermitised ground Zeeman block, eigenvalue i paired
                        with bare energy i, d_sq_eig = |U^+ |d||^2, two passes
                        over beams, 12 x 12 generator with its last row
                        replaced by ones, LU solve, clamp and normalise
* ``leapfrog``          MOTSimulator3D.simulate_trajectory_3d (:439-502)

``eig`` may be passed to force() to replace numpy's eigh (a function
H -> (w, U)); used by the tests to quantify the basis dependence of the
reference when eigenvalues are degenerate.

``basis="spin"`` (force, leapfrog): the basis convention of the GPU kernels
for ground Zeeman blocks that are s times spin-1/2 operators
(``spin_tables``): U = R(B) W, W = eigh(Zz), R the spin rotation taking z
to B / |B| about z x B (about y at B / |B| = -z), eigenvalues |B| lam0; the
identity at B = 0, as eigh returns for the zero matrix.  Pinned by: U is
unitary and U^H H_Z U = diag(|B| lam0) (tests/test_oracle_sim3d.py); on the
+z axis, where U = W, the force equals the reference's own outputs
(tests/golden/ref_force3d_axis.npz) to the precision the reference's
eigensolver has inside the near-degenerate eigenspaces of the SrOH tables.

Pinned by (tests/test_oracle_sim3d.py): the two-level saturated closed form
for one beam along B; the squared Wigner d^1 weights at theta = 0, pi/2, pi;
F(0) ~ 0 and the restoring F_z of the reference tests
(tests/test_all_functions.py:534-555); M p = 0 and sum p = 1; leapfrog on a
constant force against the closed form.
"""
import numpy as np
from scipy.linalg import solve as _lu_solve

HBAR = 1.054571817e-34
_DIRS = (np.array([1.0, 0.0, 0.0]), np.array([-1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0]),
         np.array([0.0, -1.0, 0.0]), np.array([0.0, 0.0, 1.0]), np.array([0.0, 0.0, -1.0]))


def six_beams(beam_pairs, beam_radius=10e-3):
    """List of dicts {k_hat, freq, q, s0, w}."""
    out = []
    for bp in beam_pairs:
        for i, d in enumerate(_DIRS):
            out.append(dict(k_hat=d.copy(), freq=bp["freq"], q=bp["q_fwd"] if i % 2 == 0 else bp["q_bwd"],
                            s0=bp["s0"], w=beam_radius))
    return out


def pol_weights(q_lab, B_hat):
    n = np.linalg.norm(B_hat)
    if n < 1e-15:
        return np.array([1.0, 1.0, 1.0]) / 3.0
    B_hat = B_hat / n
    ct = np.clip(B_hat[2], -1.0, 1.0)
    st = np.sin(np.arccos(ct))
    d = np.array([[((1 + ct) / 2) ** 2, (st / np.sqrt(2)) ** 2, ((1 - ct) / 2) ** 2],
                  [(st / np.sqrt(2)) ** 2, ct ** 2, (st / np.sqrt(2)) ** 2],
                  [((1 - ct) / 2) ** 2, (st / np.sqrt(2)) ** 2, ((1 + ct) / 2) ** 2]])
    w = d[:, q_lab].copy()
    s = np.sum(w)
    if s > 1e-15:
        w /= s
    return w


def zeeman_block(mol, B):
    ng = mol.n_ground
    H = B[0] * mol.zeeman_x[:ng, :ng] + B[1] * mol.zeeman_y[:ng, :ng] + B[2] * mol.zeeman_z[:ng, :ng]
    return 0.5 * (H + H.conj().T)


def spin_tables(mol, tol=1e-9):
    """(s, lam0, W, Zx, Zy) of the Hermitised ground Zeeman blocks when they
    are s times spin-1/2 operators to relative tol ([Zx, Zy] = i s Zz and
    cyclic, Zx^2 + Zy^2 + Zz^2 = 3 s^2 / 4), else None."""
    ng = mol.n_ground
    Z = []
    for a in "xyz":
        z = np.asarray(getattr(mol, f"zeeman_{a}"))[:ng, :ng].astype(complex)
        Z.append(0.5 * (z + z.conj().T))
    Zx, Zy, Zz = Z
    C = Zx @ Zx + Zy @ Zy + Zz @ Zz
    s2 = float(np.real(np.trace(C))) / (0.75 * ng)
    if not s2 > 0:
        return None
    s = np.sqrt(s2)
    dev = max(np.max(np.abs(C - 0.75 * s2 * np.eye(ng))), np.max(np.abs(Zx @ Zy - Zy @ Zx - 1j * s * Zz)),
              np.max(np.abs(Zy @ Zz - Zz @ Zy - 1j * s * Zx)), np.max(np.abs(Zz @ Zx - Zx @ Zz - 1j * s * Zy)))
    if not dev <= tol * s2:
        return None
    lam0, W = np.linalg.eigh(Zz)
    return s, lam0, W, Zx, Zy


def spin_eig(tabs, B):
    """(w, U) of the spin-rotation convention at field B (G)."""
    s, lam0, W, Zx, Zy = tabs
    ng = len(lam0)
    Bm = float(np.linalg.norm(B))
    if not Bm > 0:
        return np.zeros(ng), np.eye(ng, dtype=complex)
    n = np.asarray(B, float) / Bm if Bm > 1e-10 else np.array([0.0, 0.0, 1.0])
    c = np.sqrt(max(0.0, 0.5 * (1.0 + n[2])))
    if c > 1e-8:
        al, be = -n[1] / (2 * c), n[0] / (2 * c)
    else:
        c, al, be = 0.0, 0.0, 1.0
    Rh = c * np.eye(ng) + 1j * (al * (2.0 / s) * Zx + be * (2.0 / s) * Zy)   # R^H
    return Bm * lam0, Rh.conj().T @ W


def force(mol, beams, r, v, B_gradient, Gamma_eff_factor=1.0, hbar=HBAR, eig=None, basis="eigh"):
    """(F (3,) N, R rad/s, p (n_g,)) at one point (simulator_3d.py:168-365)."""
    r = np.asarray(r, float)
    v = np.asarray(v, float)
    G, k = mol.Gamma, mol.k
    ng, ne = mol.n_ground, mol.n_excited
    B = np.array([-B_gradient * r[0] / 2.0 * 1e4, -B_gradient * r[1] / 2.0 * 1e4, B_gradient * r[2] * 1e4])
    Bm = np.linalg.norm(B)
    B_hat = B / Bm if Bm > 1e-10 else np.array([0.0, 0.0, 1.0])
    H = zeeman_block(mol, B)
    if basis == "spin":
        tabs = spin_tables(mol)
        if tabs is None:
            raise ValueError("basis='spin' needs spin-1/2 ground Zeeman blocks")
        w, U = spin_eig(tabs, B)
    else:
        w, U = (eig or np.linalg.eigh)(H)
    Es_g = mol.energies[:ng] + w * G / (2 * np.pi)
    Es_e = mol.energies[ng:].copy()
    Es_e += B[2] * np.real(np.diag(mol.zeeman_z[ng:, ng:])) * G / (2 * np.pi)
    dsq = np.zeros((ng, ne, 3))
    for ie in range(ne):
        for q in range(3):
            d_bare = np.array([abs(mol.tdm[ig, ng + ie, q]) for ig in range(ng)])
            dsq[:, ie, q] = np.abs(U.conj().T @ d_bare) ** 2
    BR = np.zeros((ne, ng))
    for ie in range(ne):
        tot = np.sum(dsq[:, ie, :])
        if tot > 1e-30:
            for ig in range(ng):
                BR[ie, ig] = np.sum(dsq[ig, ie, :]) / tot
    gh = G / 2.0
    gh2 = gh ** 2
    s_total = np.zeros(ng)
    rates = []
    for b in beams:
        dop = -k * np.dot(b["k_hat"], v) / (2 * np.pi)
        if b["w"] is not None:
            rpar = np.dot(r, b["k_hat"])
            rp2 = max(np.dot(r, r) - rpar ** 2, 0.0)
            gauss = np.exp(-2.0 * rp2 / b["w"] ** 2)
        else:
            gauss = 1.0
        s_eff = b["s0"] * gauss
        pw = pol_weights(b["q"], B_hat)
        Rb = np.zeros((ng, ne))
        for ig in range(ng):
            for ie in range(ne):
                om = Es_e[ie] - Es_g[ig]
                rs = 0.0
                for ql in range(3):
                    d2 = dsq[ig, ie, ql]
                    if d2 < 1e-15:
                        continue
                    de = (b["freq"] + dop - om) * 2 * np.pi
                    rs += pw[ql] * d2 * (gh2 / (de ** 2 + gh2))
                s_total[ig] += s_eff * rs
                Rb[ig, ie] = rs * s_eff
        rates.append(Rb)
    Rt = np.zeros((ng, ne))
    for Rb in rates:
        for ig in range(ng):
            for ie in range(ne):
                Rt[ig, ie] += gh * Rb[ig, ie] * Gamma_eff_factor / (1.0 + s_total[ig])
    M = np.zeros((ng, ng))
    for ig in range(ng):
        M[ig, ig] -= np.sum(Rt[ig, :])
        for ik in range(ng):
            for ie in range(ne):
                M[ig, ik] += Rt[ik, ie] * BR[ie, ig]
    Ms = M.copy()
    Ms[-1, :] = 1.0
    rhs = np.zeros(ng)
    rhs[-1] = 1.0
    try:
        p = _lu_solve(Ms, rhs)
    except Exception:
        p = np.ones(ng) / ng
    p = np.maximum(p, 0.0)
    ps = np.sum(p)
    p = p / ps if ps > 0 else np.ones(ng) / ng
    F = np.zeros(3)
    R = 0.0
    for b, Rb in zip(beams, rates):
        for ig in range(ng):
            for ie in range(ne):
                rate = gh * Rb[ig, ie] * Gamma_eff_factor / (1.0 + s_total[ig])
                F += hbar * k * b["k_hat"] * rate * p[ig]
                R += rate * p[ig]
    return F, R, p, M


def leapfrog(mol, beams, r0, v0, t_max, dt, B_gradient, Gamma_eff_factor=1.0, r_escape=0.015, hbar=HBAR,
             n_steps=None, force_fn=None, basis="eigh"):
    """(t, r (n,3), v (n,3)) of simulate_trajectory_3d (simulator_3d.py:439-502)."""
    m = mol.mass
    n = int(t_max / dt) if n_steps is None else n_steps
    f = force_fn or (lambda rr, vv: force(mol, beams, rr, vv, B_gradient, Gamma_eff_factor, hbar,
                                          basis=basis)[0])
    t = np.arange(n) * dt
    r = np.zeros((n, 3))
    v = np.zeros((n, 3))
    r[0] = r0
    v[0] = v0
    vh = v[0] + 0.5 * dt * f(r[0], v[0]) / m
    for j in range(1, n):
        r[j] = r[j - 1] + dt * vh
        if np.linalg.norm(r[j]) > r_escape:
            r[j:] = r[j]
            v[j:] = vh
            break
        F = f(r[j], vh)
        vn = vh + dt * F / m
        v[j] = 0.5 * (vh + vn)
        vh = vn
    return t, r, v
