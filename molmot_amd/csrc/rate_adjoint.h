// Reverse-mode gradient of the rate-equation force at one (v, z) point.
//
// The reference has no derivatives beyond two-point finite differences
// (molmot/mot/force_scan.py:84-128).  This computes, for one point, F and
//     dF/d(detuning_c)  c < NC   (Hz offset added to every beam of component c)
//     dF/d(split)                (beam c moves by split_sign_c * split, Hz)
//     dF/d(s0 scale)             (every beam's s0 multiplied by lambda, at lambda = 1)
//     dF/d(B')                   (T/m)
//     dF/dz, dF/dv               (m, m/s)
// by one forward pass of the reduced steady-state solve (rate_core.h) and
// one reverse pass.  The stationary vector y of the excited-state chain
// (G y = 0, y_0 = 1, G the generator built from K) is differentiated
// implicitly: with G^ = G whose row 0 is replaced by e_0^T, lambda solves
// G^T lambda = ybar and Gbar_ij = -lambda_i y_j (i > 0), so the GTH
// elimination itself is not stepped back through.
#pragma once

#include "rate_core.h"

#define MM_ADJ_NC 4                  // frequency components with detuning gradients
#define MM_ADJ_NJ (MM_ADJ_NC + 5)    // detunings, split, s0 scale, B', z, v
#define MM_ADJ_J_SPLIT MM_ADJ_NC
#define MM_ADJ_J_S0 (MM_ADJ_NC + 1)
#define MM_ADJ_J_B (MM_ADJ_NC + 2)
#define MM_ADJ_J_Z (MM_ADJ_NC + 3)
#define MM_ADJ_J_V (MM_ADJ_NC + 4)

// Per-thread LDS slab layout (stride ls between consecutive doubles):
//   [0, NG*NE)     s_ge
//   [NG*NE + g]    T_g
// S_g, D_g and Fg_g are recomputed from these where needed.
template <int NG, int NE>
constexpr int mm_adj_slab() { return NG * NE + NG; }

template <int NG, int NE, bool UNIT>
__device__ __forceinline__ bool rate_grad_fast(const MMRatePlanDev* __restrict__ P, double v, double z,
                                               double Bgrad, double& F, double (&J)[MM_ADJ_NJ], double* slab,
                                               int ls) {
    constexpr int O_T = NG * NE;
    const double c1 = P->gh_geff, c2 = P->fk_gh_geff;
    double Ee[NE];
    const double Bg = Bgrad * z * 1e4;
    const double zfac = P->Gamma / MM_TWO_PI;  // dEs/d(Bg*zd)
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const double zdi = P->zd[NG + i];
        Ee[i] = (zdi == 0.0) ? P->E[NG + i] : P->E[NG + i] + Bg * zdi * P->Gamma / MM_TWO_PI;
    }
    const double dop = UNIT ? P->mk * v / MM_TWO_PI : 0.0;

    // ---------------- forward, keeping what the reverse pass needs ----------------
    double K[NE][NE];
#pragma unroll
    for (int e = 0; e < NE; e++)
#pragma unroll
        for (int f = 0; f < NE; f++) K[e][f] = 0.0;
    bool ok = true;
    const bool packed = UNIT && NE <= 4 && P->packed && NE * NG <= MM_MAX_G * 4;
    auto brf = [&](int f, int g) -> double { return packed ? P->grec[g].br[f & 3] : P->BR[f][g]; };
    auto Egof = [&](int g) -> double {
        const double zdg = packed ? P->grec[g].zd : P->zd[g];
        const double E0 = packed ? P->grec[g].E : P->E[g];
        return (zdg == 0.0) ? E0 : E0 + Bg * zdg * P->Gamma / MM_TWO_PI;
    };
#pragma unroll 1
    for (int g = 0; g < NG; g++) {
        const double Eg = Egof(g);
        double S[NE];
        double Sg = 0.0, Tg = 0.0;
        const unsigned mask = packed ? P->pair_mask[g] : 0u;
#pragma unroll
        for (int e = 0; e < NE; e++) {
            double s = 0.0, t = 0.0;
            if (packed) {
                if (mask & (1u << e)) mm_pair4(P->prec[g * 4 + (e & 3)], P->gq, Ee[e] - Eg, dop, s, t);
            } else {
                mm_pair_sums<UNIT>(P, g * NE + e, Ee[e] - Eg, v, dop, s, t);
            }
            S[e] = s;
            slab[(g * NE + e) * ls] = s;
            Sg += s;
            Tg += t;
        }
        ok = ok && (Sg > 0.0);
        const double rS = mm_rcp(Sg);
#pragma unroll
        for (int e = 0; e < NE; e++) {
            const double a = S[e] * rS;
#pragma unroll
            for (int f = 0; f < NE; f++)
                if (f != e) K[e][f] = fma(a, brf(f, g), K[e][f]);
        }
        slab[(O_T + g) * ls] = Tg;
    }
    // generator of the excited chain from the off-diagonal K (kept for the adjoint)
    double G[NE][NE];
#pragma unroll
    for (int e = 0; e < NE; e++) {
        double out = 0.0;
#pragma unroll
        for (int f = 0; f < NE; f++) {
            G[e][f] = (f != e) ? K[e][f] : 0.0;
            if (f != e) out += K[f][e];
        }
        G[e][e] = -out;
    }
#pragma unroll
    for (int m = NE - 1; m >= 1; m--) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < m; j++) s += K[j][m];
        ok = ok && (s > 0.0);
        const double rs = mm_rcp(s);
#pragma unroll
        for (int i = 0; i < m; i++) K[m][i] *= rs;
#pragma unroll
        for (int i = 0; i < m; i++)
#pragma unroll
            for (int j = 0; j < m; j++)
                if (i != j) K[j][i] = fma(K[m][i], K[j][m], K[j][i]);
    }
    if (!ok) return false;
    double y[NE];
    y[0] = 1.0;
#pragma unroll
    for (int j = 1; j < NE; j++) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < j; i++) acc = fma(y[i], K[j][i], acc);
        y[j] = acc;
    }
    // D_g, Fg_g and q_g = u_g / D_g of ground g, from the stored sums
    // Sg, Tg, sat = 1/(1+Sg), rS = 1/Sg, rD = 1/D_g (D_g = c1 Sg sat), Fg_g and
    // u_g of ground g from the stored sums, with one reciprocal
    auto ground = [&](int g, double& Sg, double& Tg, double& sat, double& rS, double& rD, double& Fg, double& u) {
        Sg = 0.0;
#pragma unroll
        for (int e = 0; e < NE; e++) Sg += slab[(g * NE + e) * ls];
        Tg = slab[(O_T + g) * ls];
        const double s1 = 1.0 + Sg;
        const double rr = mm_rcp(Sg * s1);
        sat = rr * Sg;
        rS = rr * s1;
        rD = (rS * s1) * P->inv_gh_geff;
        Fg = c2 * sat * Tg;
        u = 0.0;
#pragma unroll
        for (int f = 0; f < NE; f++) u = fma(y[f], brf(f, g), u);
    };
    double Z = 0.0, Fs = 0.0;
#pragma unroll 1
    for (int g = 0; g < NG; g++) {
        double Sg, Tg, sat, rS, rD, Fg, u;
        ground(g, Sg, Tg, sat, rS, rD, Fg, u);
        const double q = u * rD;
        Z += q;
        Fs = fma(q, Fg, Fs);
    }
    const double rZ = mm_rcp(Z);
    F = Fs * rZ;

    // ---------------- reverse ----------------
    double yb[NE];
#pragma unroll
    for (int f = 0; f < NE; f++) yb[f] = 0.0;
#pragma unroll 1
    for (int g = 0; g < NG; g++) {
        double Sg, Tg, sat, rS, rD, Fg, u;
        ground(g, Sg, Tg, sat, rS, rD, Fg, u);
        const double ub = (Fg - F) * rZ * rD;
#pragma unroll
        for (int f = 0; f < NE; f++) yb[f] = fma(ub, brf(f, g), yb[f]);
    }
    // lambda = G^-T ybar with row 0 of G replaced by e_0^T, i.e. A = G^T with
    // column 0 replaced by e_0: partial-pivoting elimination, NE x NE
    double A[NE][NE], lam[NE];
#pragma unroll
    for (int i = 0; i < NE; i++) {
#pragma unroll
        for (int j = 0; j < NE; j++) A[i][j] = (j == 0) ? (i == 0 ? 1.0 : 0.0) : G[j][i];
        lam[i] = yb[i];
    }
#pragma unroll
    for (int c = 0; c < NE; c++) {
#pragma unroll
        for (int r = c + 1; r < NE; r++) {
            if (fabs(A[r][c]) > fabs(A[c][c])) {
#pragma unroll
                for (int j = 0; j < NE; j++) {
                    const double t = A[c][j];
                    A[c][j] = A[r][j];
                    A[r][j] = t;
                }
                const double t = lam[c];
                lam[c] = lam[r];
                lam[r] = t;
            }
        }
        const double rp = 1.0 / A[c][c];
#pragma unroll
        for (int r = c + 1; r < NE; r++) {
            const double fct = A[r][c] * rp;
#pragma unroll
            for (int j = c; j < NE; j++) A[r][j] = fma(-fct, A[c][j], A[r][j]);
            lam[r] = fma(-fct, lam[c], lam[r]);
        }
    }
#pragma unroll
    for (int i = NE - 1; i >= 0; i--) {
        double s = lam[i];
#pragma unroll
        for (int j = i + 1; j < NE; j++) s = fma(-A[i][j], lam[j], s);
        lam[i] = s / A[i][i];
    }
    // Kbar[e][f] = Gbar[e][f] - Gbar[f][f], Gbar[i][j] = -lam_i y_j for i > 0
    double Kb[NE][NE];
#pragma unroll
    for (int e = 0; e < NE; e++)
#pragma unroll
        for (int f = 0; f < NE; f++) {
            const double gef = (e > 0) ? -lam[e] * y[f] : 0.0;
            const double gff = (f > 0) ? -lam[f] * y[f] : 0.0;
            Kb[e][f] = (e != f) ? gef - gff : 0.0;
        }

    double dcomp[MM_ADJ_NC];
#pragma unroll
    for (int c = 0; c < MM_ADJ_NC; c++) dcomp[c] = 0.0;
    double dopb = 0.0, lamb = 0.0, Bgb = 0.0;
    double Eeb[NE];
#pragma unroll
    for (int e = 0; e < NE; e++) Eeb[e] = 0.0;
#pragma unroll 1
    for (int g = 0; g < NG; g++) {
        const double zdg = packed ? P->grec[g].zd : P->zd[g];
        const double Eg = Egof(g);
        double Sg, Tg, sat, rS, rD, Fg, u;
        ground(g, Sg, Tg, sat, rS, rD, Fg, u);
        const double q = u * rD;
        const double qb = (Fg - F) * rZ;
        const double Db = -qb * q * rD;   // dF/dD_g
        const double Fgb = q * rZ;        // dF/dFg_g
        double ab[NE], a[NE];
        double Sb = 0.0;
#pragma unroll
        for (int e = 0; e < NE; e++) {
            double acc = 0.0;
#pragma unroll
            for (int f = 0; f < NE; f++)
                if (f != e) acc = fma(Kb[e][f], brf(f, g), acc);
            ab[e] = acc;
            a[e] = slab[(g * NE + e) * ls] * rS;
            Sb = fma(-acc, a[e] * rS, Sb);
        }
        Sb = fma(Db * c1, sat * sat, Sb);
        Sb = fma(-Fgb * c2 * Tg, sat * sat, Sb);
        const double Tb = Fgb * c2 * sat;
        double Egb = 0.0;
        const unsigned mask = packed ? P->pair_mask[g] : 0u;
#pragma unroll
        for (int e = 0; e < NE; e++) {
            const double sb = fma(ab[e], rS, Sb);
            const double om = Ee[e] - Eg;
            double omb = 0.0;
            if (packed) {
                if (mask & (1u << e)) {
                    // four terms (+z, +z, -z, -z), reciprocals from one shared inverse
                    const MMPairRec& Q = P->prec[g * 4 + (e & 3)];
                    const unsigned pc = P->pair_comp[g * 4 + (e & 3)];
                    double d[4], x[4];
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        d[j] = (j < 2 ? Q.f[j] + dop : Q.f[j] - dop) - om;
                        x[j] = fma(d[j], d[j], P->gq);
                    }
                    const double pab = x[0] * x[1], pcd = x[2] * x[3];
                    const double rr = mm_rcp1(pab * pcd);
                    const double rp = rr * pcd, rm = rr * pab;
                    const double r[4] = {x[1] * rp, x[0] * rp, x[3] * rm, x[2] * rm};
                    const double wbp = Tb + sb, wbm = sb - Tb;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const double w = Q.c2[j] * r[j];
                        const double wb = j < 2 ? wbp : wbm;
                        lamb = fma(wb, w, lamb);
                        const double db = -2.0 * wb * w * r[j] * d[j];
                        omb -= db;
                        dopb = j < 2 ? dopb + db : dopb - db;
                        const unsigned c = (pc >> (8 * j)) & 0xffu;
                        if (c == 0) dcomp[0] += db;
                        else if (c == 1) dcomp[1] += db;
                        else if (c == 2) dcomp[2] += db;
                        else if (c == 3) dcomp[3] += db;
                    }
                }
                Eeb[e] += omb;
                Egb -= omb;
                continue;
            }
            const int t0 = P->pair_start[g * NE + e];
            const int t1 = P->pair_start[g * NE + e + 1];
            for (int it = t0; it < t1; ++it) {
                const MMTerm& T = P->terms[it];
                const double fd = UNIT ? fma(T.kdir, dop, T.freq) : T.freq + P->mk * T.kdir * v / MM_TWO_PI;
                const double d = fd - om;
                const double r = mm_rcp1(fma(d, d, P->gq));
                const double w = T.C2 * r;
                const double wb = fma(T.kdir, Tb, sb);
                lamb = fma(wb, w, lamb);
                const double db = -2.0 * wb * w * r * d;
                omb -= db;
                dopb = fma(T.kdir, db, dopb);
                const int c = T.comp;
                if (c == 0) dcomp[0] += db;
                else if (c == 1) dcomp[1] += db;
                else if (c == 2) dcomp[2] += db;
                else if (c == 3) dcomp[3] += db;
            }
            Eeb[e] += omb;
            Egb -= omb;
        }
        Bgb = fma(Egb, zdg * zfac, Bgb);
    }
#pragma unroll
    for (int e = 0; e < NE; e++) Bgb = fma(Eeb[e], P->zd[NG + e] * zfac, Bgb);
    double split = 0.0;
#pragma unroll
    for (int c = 0; c < MM_ADJ_NC; c++) {
        J[c] = dcomp[c];
        split = fma(P->split_sign[c], dcomp[c], split);
    }
    J[MM_ADJ_J_SPLIT] = split;
    J[MM_ADJ_J_S0] = lamb;
    J[MM_ADJ_J_B] = Bgb * z * 1e4;
    J[MM_ADJ_J_Z] = Bgb * Bgrad * 1e4;
    J[MM_ADJ_J_V] = dopb * P->mk / MM_TWO_PI;
    return true;
}
