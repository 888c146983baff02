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
//   [g]        rS_g = 1/S_g
//   [NG + g]   B_g  = T_g / (1 + S_g)
// The reverse pass re-evaluates the pair terms of each ground state and
// defers the part of every term adjoint that is proportional to Sbar_g (the
// adjoint of S_g, known only once all pairs of g have been seen): the
// S-weighted sums are collected per component, per pair and per beam
// direction and scaled by Sbar_g at the end of the ground state.  The slab
// is therefore the same size as the force kernel's, which keeps the LDS
// footprint from limiting occupancy.
template <int NG, int NE, int SL = 2>
constexpr int mm_adj_slab() { return SL * NG; }

// GU1..GU4: unroll factors of the four ground-state loops (forward pairs,
// force sums, ybar, reverse pairs); 1 keeps the register footprint small.
// CO: the plan has component-ordered pair records (comp_order; implies
// UNIT and packed); the reverse pass then reads them, otherwise it walks
// the general term list.  A compile-time choice, so that neither pass
// carries the other's code and its register moves.
// SL: slab values per ground state, 2 or 3; with 3 the forward pass also
// keeps 1/(1 + S_g), which the reverse pass then reads instead of
// recomputing (for kernels whose occupancy the larger slab does not limit).
template <int NG, int NE, bool UNIT, int GU1 = 1, int GU2 = 1, int GU3 = 1, int GU4 = 1, bool CO = false,
          int SL = 2>
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
        Ee[i] = (zdi == 0.0) ? P->E[NG + i] : P->E[NG + i] + mm_div_rn(Bg * zdi * P->Gamma, MM_TWO_PI, MM_INV_TWO_PI);
    }
    const double dop = UNIT ? mm_div_rn(P->mk * v, MM_TWO_PI, MM_INV_TWO_PI) : 0.0;

    // ---------------- forward, keeping what the reverse pass needs ----------------
    double K[NE][NE];
#pragma unroll
    for (int e = 0; e < NE; e++)
#pragma unroll
        for (int f = 0; f < NE; f++) K[e][f] = 0.0;
    bool ok = true;
    static_assert(!CO || (UNIT && NE <= 4 && MM_ADJ_NC == 4), "component-ordered records need unit directions");
    const bool packed = CO || (UNIT && NE <= 4 && P->packed && NE * NG <= MM_MAX_G * 4);
    const double gq = P->gq;
    // per-ground values: from the 64-byte ground record in the CO variant
    // (one scalar load), else from the plain tables (no run-time choice
    // between two addresses per load)
    auto brf = [&](int f, int g) -> double {
        if constexpr (CO) return P->grec[g].br[f & 3];
        else return P->BR[f][g];
    };
    auto Egof = [&](int g) -> double {
        if constexpr (CO) {
            const MMGroundRec& G = P->grec[g];
            return (G.zd == 0.0) ? G.E : G.E + mm_div_rn(Bg * G.zd * G.Gamma, MM_TWO_PI, MM_INV_TWO_PI);
        } else {
            const double zdg = P->zd[g];
            const double E0 = P->E[g];
            return (zdg == 0.0) ? E0 : E0 + mm_div_rn(Bg * zdg * P->Gamma, MM_TWO_PI, MM_INV_TWO_PI);
        }
    };
#pragma unroll GU1
    for (int g = 0; g < NG; g++) {
        const double Eg = Egof(g);
        double S[NE];
        double Sg = 0.0, Tg = 0.0;
        const unsigned mask = CO ? P->grec[g].mask : packed ? P->pair_mask[g] : 0u;
#pragma unroll
        for (int e = 0; e < NE; e++) {
            double s = 0.0, t = 0.0;
            S[e] = 0.0;
            if (packed) {
                // empty pairs add nothing (the sums stay bit-identical)
                if (mask & (1u << e)) {
                    mm_pair4(P->prec[g * 4 + (e & 3)], gq, Ee[e] - Eg, dop, s, t);
                    S[e] = s;
                    Sg += s;
                    Tg += t;
                }
            } else {
                mm_pair_sums<UNIT>(P, g * NE + e, Ee[e] - Eg, v, dop, s, t);
                S[e] = s;
                Sg += s;
                Tg += t;
            }
        }
        ok = ok && (Sg > 0.0);
        // rS = 1/S_g and sat = 1/(1 + S_g) from one reciprocal
        const double s1 = 1.0 + Sg;
        const double rr = mm_rcp(Sg * s1);
        const double rS = rr * s1;
#pragma unroll
        for (int e = 0; e < NE; e++) {
            const double a = S[e] * rS;
#pragma unroll
            for (int f = 0; f < NE; f++)
                if (f != e) K[e][f] = fma(a, brf(f, g), K[e][f]);
        }
        slab[g * ls] = rS;
        slab[(NG + g) * ls] = (rr * Sg) * Tg;
        if constexpr (SL == 3) slab[(2 * NG + g) * ls] = rr * Sg;
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
    // 1/D_g = (1 + 1/S_g) / (Gamma/2 Geff), Fg_g = c2 T_g/(1 + S_g), u_g = sum_f y_f br_fg
    auto uof = [&](int g) -> double {
        double u = 0.0;
#pragma unroll
        for (int f = 0; f < NE; f++) u = fma(y[f], brf(f, g), u);
        return u;
    };
    double Z = 0.0, Fs = 0.0;
#pragma unroll GU2
    for (int g = 0; g < NG; g++) {
        const double rD = (1.0 + slab[g * ls]) * P->inv_gh_geff;
        const double q = uof(g) * rD;
        Z += q;
        Fs = fma(q, c2 * slab[(NG + g) * ls], Fs);
    }
    const double rZ = mm_rcp(Z);
    F = Fs * rZ;

    // ---------------- reverse ----------------
    double yb[NE];
#pragma unroll
    for (int f = 0; f < NE; f++) yb[f] = 0.0;
#pragma unroll GU3
    for (int g = 0; g < NG; g++) {
        const double rD = (1.0 + slab[g * ls]) * P->inv_gh_geff;
        const double ub = (c2 * slab[(NG + g) * ls] - F) * rZ * rD;
#pragma unroll
        for (int f = 0; f < NE; f++) yb[f] = fma(ub, brf(f, g), yb[f]);
    }
    // lambda = G^-T ybar with row 0 of G replaced by e_0^T, i.e. A = G^T with
    // column 0 replaced by e_0.  Column 0 needs no elimination, and the
    // trailing block (G^T)[1:, 1:] is diagonally dominant by rows (a column
    // of the generator sums to zero and its off-diagonal entries are >= 0),
    // so Gaussian elimination without pivoting is stable there (growth
    // factor <= 2) and every pivot is negative.
    double A[NE][NE], lam[NE], rdg[NE];
#pragma unroll
    for (int i = 0; i < NE; i++) {
#pragma unroll
        for (int j = 0; j < NE; j++) A[i][j] = (j == 0) ? (i == 0 ? 1.0 : 0.0) : G[j][i];
        lam[i] = yb[i];
    }
    rdg[0] = 1.0;
#pragma unroll
    for (int c = 1; c < NE; c++) {
        rdg[c] = mm_rcp1(A[c][c]);
#pragma unroll
        for (int r = c + 1; r < NE; r++) {
            const double fct = A[r][c] * rdg[c];
#pragma unroll
            for (int j = c + 1; j < NE; j++) A[r][j] = fma(-fct, A[c][j], A[r][j]);
            lam[r] = fma(-fct, lam[c], lam[r]);
        }
    }
#pragma unroll
    for (int i = NE - 1; i >= 0; i--) {
        double s = lam[i];
#pragma unroll
        for (int j = i + 1; j < NE; j++) s = fma(-A[i][j], lam[j], s);
        lam[i] = s * rdg[i];
    }
    // Kbar[e][f] = Gbar[e][f] - Gbar[f][f] with Gbar[i][j] = -lam_i y_j for
    // i > 0, so the adjoint of the branching sum of pair (g, e),
    // ab_e = sum_{f != e} Kbar[e][f] br_fg, reduces to L_g - lam'_e u_g with
    // u_g = sum_f y_f br_fg, L_g = sum_{f > 0} lam_f y_f br_fg and lam'_0 = 0
    double lamp[NE], ly[NE];
#pragma unroll
    for (int e = 0; e < NE; e++) {
        lamp[e] = e > 0 ? lam[e] : 0.0;
        ly[e] = e > 0 ? lam[e] * y[e] : 0.0;
    }

    // Term adjoint: with D = w r d (w = C2 r, r = 1/(d^2 + gq)) and the
    // weight wb = v + Sbar_g, v = ab_e rS +- Tbar_g, the detuning adjoint is
    // db = -2 v D - 2 Sbar_g D.  The first part is accumulated at once; the
    // second through the per-g sums cS (per component), pS (per pair) and
    // kS (signed by beam direction), scaled by Sbar_g after the last pair.
    double dcomp[MM_ADJ_NC];
#pragma unroll
    for (int c = 0; c < MM_ADJ_NC; c++) dcomp[c] = 0.0;
    double dopb = 0.0, lamb = 0.0, Bgb = 0.0;
    double Eeb[NE], zez[NE];
#pragma unroll
    for (int e = 0; e < NE; e++) {
        Eeb[e] = 0.0;
        zez[e] = P->zd[NG + e] * zfac;
    }
#pragma unroll GU4
    for (int g = 0; g < NG; g++) {
        const double zdg = CO ? P->grec[g].zd : P->zd[g];
        const double Eg = Egof(g);
        const double rS = slab[g * ls];
        const double Bsat = slab[(NG + g) * ls];
        double sat;
        if constexpr (SL == 3) sat = slab[(2 * NG + g) * ls];
        else sat = rS * mm_rcp1(1.0 + rS);  // reverse-only: ~1e-15 is ample
        const double rD = (1.0 + rS) * P->inv_gh_geff;
        double u = 0.0, L = 0.0;
#pragma unroll
        for (int f = 0; f < NE; f++) {
            const double b = brf(f, g);
            u = fma(y[f], b, u);
            L = fma(ly[f], b, L);
        }
        const double q = u * rD;
        const double qb = (c2 * Bsat - F) * rZ;
        const double Db = -qb * q * rD;   // dF/dD_g
        const double Fgb = q * rZ;        // dF/dFg_g
        const double Tb = Fgb * c2 * sat;
        // the part of Sbar_g known before the pairs: D_g and Fg_g through sat
        double Sb = Db * c1 * sat * sat - Fgb * c2 * Bsat * sat;
        double cS[MM_ADJ_NC], pS[NE], kS = 0.0, Sg = 0.0, Egb = 0.0;
#pragma unroll
        for (int c = 0; c < MM_ADJ_NC; c++) cS[c] = 0.0;
        if constexpr (CO) {
            // only the non-empty pairs contribute.  The energy adjoint of
            // pair (g, e) reaches B' only through dom/dBg = (zd_e - zd_g)
            // zfac, so it goes straight into Bgb (PZ: the deferred part)
            const unsigned mask = P->grec[g].mask;
            const unsigned dirm = P->grec[g].dirm;
            const double zgz = zdg * zfac;
            double PZ = 0.0;
#pragma unroll
            for (int e = 0; e < NE; e++) {
                if (!(mask & (1u << e))) continue;
                const double vb = fma(-lamp[e], u, L) * rS;
                const double om = Ee[e] - Eg;
                // four terms, slot c of frequency component c, direction
                // kd = +-1 (kd dop is exact, so f + kd dop rounds as f +- dop
                // does); reciprocals from one shared inverse
                const MMPairRec& Q = P->precc[g * 4 + (e & 3)];
                const unsigned dm = dirm >> (4 * (e & 3));
                double kd[4], d[4], x[4];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    kd[j] = ((dm >> j) & 1u) ? -1.0 : 1.0;
                    d[j] = (Q.f[j] + kd[j] * dop) - om;
                    x[j] = fma(d[j], d[j], gq);
                }
                const double p01 = x[0] * x[1], p23 = x[2] * x[3];
                const double rr = mm_rcp1(p01 * p23);
                const double ra = rr * p23, rb = rr * p01;
                const double r[4] = {x[1] * ra, x[0] * ra, x[3] * rb, x[2] * rb};
                // per term, with a = v_j D: the detuning adjoint db = -2 a
                // goes to its component; omb = -sum db = 2 sum a and the
                // Doppler adjoint -2 sum kd a are formed once per pair
                double s_e = 0.0, p_e = 0.0, A = 0.0, Ak = 0.0;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const double w = Q.c2[j] * r[j];
                    const double vj = fma(kd[j], Tb, vb);
                    const double D = w * r[j] * d[j];
                    const double a = vj * D;
                    s_e = j ? s_e + w : w;
                    p_e = j ? p_e + D : D;
                    A = j ? A + a : a;
                    Ak = j ? fma(kd[j], a, Ak) : kd[0] * a;
                    lamb = fma(vj, w, lamb);
                    kS = fma(kd[j], D, kS);
                    dcomp[j] = fma(-2.0, a, dcomp[j]);
                    cS[j] += D;
                }
                dopb = fma(-2.0, Ak, dopb);
                // Sbar_g collects -ab_e s_e rS^2 from every pair
                Sb = fma(-vb * rS, s_e, Sb);
                Sg += s_e;
                const double dz = zez[e] - zgz;
                Bgb = fma(2.0 * A, dz, Bgb);
                PZ = fma(p_e, dz, PZ);
            }
            // deferred Sbar_g parts: db = -2 Sbar_g D per term, lambda-bar gets Sbar_g S_g
            const double m2 = -2.0 * Sb;
            lamb = fma(Sb, Sg, lamb);
            dopb = fma(m2, kS, dopb);
#pragma unroll
            for (int c = 0; c < MM_ADJ_NC; c++) dcomp[c] = fma(m2, cS[c], dcomp[c]);
            Bgb = fma(-m2, PZ, Bgb);
        } else {
#pragma unroll
            for (int e = 0; e < NE; e++) {
                const double vb = fma(-lamp[e], u, L) * rS;
                const double om = Ee[e] - Eg;
                double omb = 0.0, s_e = 0.0, p_e = 0.0;
                const int t0 = P->pair_start[g * NE + e];
                const int t1 = P->pair_start[g * NE + e + 1];
                for (int it = t0; it < t1; ++it) {
                    const MMTerm& T = P->terms[it];
                    const double fd = UNIT ? fma(T.kdir, dop, T.freq) : T.freq + mm_div_rn(P->mk * T.kdir * v, MM_TWO_PI, MM_INV_TWO_PI);
                    const double d = fd - om;
                    const double r = mm_rcp1(fma(d, d, gq));
                    const double w = T.C2 * r;
                    const double vj = fma(T.kdir, Tb, vb);
                    s_e += w;
                    lamb = fma(vj, w, lamb);
                    const double D = w * r * d;
                    const double db = -2.0 * vj * D;
                    p_e += D;
                    omb -= db;
                    dopb = fma(T.kdir, db, dopb);
                    kS = fma(T.kdir, D, kS);
                    const int c = T.comp;
                    if (c == 0) {
                        dcomp[0] += db;
                        cS[0] += D;
                    } else if (c == 1) {
                        dcomp[1] += db;
                        cS[1] += D;
                    } else if (c == 2) {
                        dcomp[2] += db;
                        cS[2] += D;
                    } else if (c == 3) {
                        dcomp[3] += db;
                        cS[3] += D;
                    }
                }
                // Sbar_g collects -ab_e s_e rS^2 from every pair
                Sb = fma(-vb * rS, s_e, Sb);
                Sg += s_e;
                pS[e] = p_e;
                Eeb[e] += omb;
                Egb -= omb;
            }
            // deferred Sbar_g parts: db = -2 Sbar_g D per term, lambda-bar gets Sbar_g S_g
            const double m2 = -2.0 * Sb;
            lamb = fma(Sb, Sg, lamb);
            dopb = fma(m2, kS, dopb);
#pragma unroll
            for (int c = 0; c < MM_ADJ_NC; c++) dcomp[c] = fma(m2, cS[c], dcomp[c]);
#pragma unroll
            for (int e = 0; e < NE; e++) {
                const double omS = -m2 * pS[e];
                Eeb[e] += omS;
                Egb -= omS;
            }
            Bgb = fma(Egb, zdg * zfac, Bgb);
        }
    }
    if constexpr (!CO) {
#pragma unroll
        for (int e = 0; e < NE; e++) Bgb = fma(Eeb[e], zez[e], Bgb);
    }
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
