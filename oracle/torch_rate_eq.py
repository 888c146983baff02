"""Torch FP64 (CPU) restatement of the reference rate-equation force,
differentiated by torch autograd.  TEST INFRASTRUCTURE ONLY: the independent
check of the gfx950 adjoint kernel.

Follows molmot/obe/rate_equations_jit.py:18-145 (one pass, saturation applied
after the beam loop) with the same detuning arithmetic
(:29-35, :63, :70-72), the population matrix of :85-97 with its last row
replaced by ones, and a dense solve (torch.linalg.solve) in place of the
hand-written elimination of :99-126 (the two agree to rounding for
non-singular systems; the oracle's tests pin that).  Parameters follow
molmot_amd.adjoint: theta = [det0..det3 (Hz), split (Hz), s0 scale, B' (T/m)].
"""
import numpy as np
import torch

TWO_PI = 2.0 * 3.141592653589793


def force(mol, bt, theta, v, z, Gamma_eff=1.0, hbar=1.054571817e-34):
    """theta: (n, 7) or (7,) float64 tensor; v, z: (n,) -> F (n,)."""
    ng, ne = mol.n_ground, mol.n_excited
    G = float(mol.Gamma)
    k = float(mol.k)
    gh = G / 2.0
    ghs = gh * gh
    n = v.shape[0]
    if theta.dim() == 1:
        theta = theta.unsqueeze(0).expand(n, -1)
    E = torch.as_tensor(mol.energies, dtype=torch.float64)
    zd = torch.as_tensor(mol.zeeman_z_diag, dtype=torch.float64)
    d2 = np.asarray(mol.d_squared)
    # branching ratios, JIT summation order
    BR = np.zeros((ne, ng))
    for ie in range(ne):
        tot = 0.0
        for ig in range(ng):
            for q in range(3):
                tot += d2[ig, ie, q]
        if tot > 1e-30:
            for ig in range(ng):
                s = 0.0
                for q in range(3):
                    s += d2[ig, ie, q]
                BR[ie, ig] = s / tot
    BR = torch.as_tensor(BR)
    Bg = theta[:, 6] * z * 1e4
    Es = E.unsqueeze(0) + Bg.unsqueeze(1) * zd.unsqueeze(0) * G / TWO_PI     # (n, ns)
    R_sum = [[torch.zeros(n, dtype=torch.float64) for _ in range(ne)] for _ in range(ng)]
    F_tot = [[torch.zeros(n, dtype=torch.float64) for _ in range(ne)] for _ in range(ng)]
    s_tot = [torch.zeros(n, dtype=torch.float64) for _ in range(ng)]
    for b in range(bt.n_beams):
        kdir = float(bt.dirs[b])
        pol = int(bt.pols[b])
        freq = float(bt.freqs[b]) + theta[:, int(bt.comp[b])] + float(bt.split_sign[b]) * theta[:, 4]
        s0 = float(bt.s0s[b]) * theta[:, 5]
        doppler = -k * kdir * v / TWO_PI
        for ig in range(ng):
            for ie in range(ne):
                dd = float(d2[ig, ie, pol])
                if dd < 1e-15:
                    continue
                om = Es[:, ng + ie] - Es[:, ig]
                delta = (freq + doppler - om) * TWO_PI
                L = ghs / (delta * delta + ghs)
                rate = gh * s0 * dd * L * Gamma_eff
                R_sum[ig][ie] = R_sum[ig][ie] + rate
                F_tot[ig][ie] = F_tot[ig][ie] + hbar * k * kdir * rate
                s_tot[ig] = s_tot[ig] + s0 * dd * L
    R = torch.stack([torch.stack([R_sum[g][e] / (1.0 + s_tot[g]) for e in range(ne)], -1) for g in range(ng)], 1)
    Ft = torch.stack([torch.stack([F_tot[g][e] / (1.0 + s_tot[g]) for e in range(ne)], -1) for g in range(ng)], 1)
    # M[ig, ik] = -delta_{ig,ik} sum_e R[ig,e] + sum_e R[ik,e] BR[e,ig]
    M = torch.einsum("nke,eg->ngk", R, BR) - torch.diag_embed(R.sum(-1))
    M = torch.cat([M[:, :-1, :], torch.ones(n, 1, ng, dtype=torch.float64)], 1)
    rhs = torch.zeros(n, ng, 1, dtype=torch.float64)
    rhs[:, -1, 0] = 1.0
    p = torch.linalg.solve(M, rhs)[..., 0]
    p = p / p.sum(-1, keepdim=True)
    return (p * Ft.sum(-1)).sum(-1)


def jacobian(mol, bt, theta0, v, z, **kw):
    """Per-point [dF/dtheta (7), dF/dz, dF/dv] -> (n, 9) numpy."""
    v = torch.as_tensor(np.asarray(v, dtype=np.float64)).clone().requires_grad_(True)
    z = torch.as_tensor(np.asarray(z, dtype=np.float64)).clone().requires_grad_(True)
    th = torch.as_tensor(np.asarray(theta0, dtype=np.float64)).unsqueeze(0).repeat(v.shape[0], 1)
    th.requires_grad_(True)
    F = force(mol, bt, th, v, z, **kw)
    gth, gz, gv = torch.autograd.grad(F.sum(), [th, z, v])
    return np.concatenate([gth.numpy(), gz.numpy()[:, None], gv.numpy()[:, None]], 1), F.detach().numpy()
