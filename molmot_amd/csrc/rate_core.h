// FP64 rate-equation evaluation for one (v, z) point, one thread per point.
//
// Reference algorithm: molmot/obe/rate_equations_jit.py:18-145 (_rate_eq_core)
// and molmot/obe/rate_equations.py:23-191.  Two stages are kept operation for
// operation because they carry catastrophic cancellation: the Zeeman-shifted
// energies Es (rate_equations_jit.py:29-35) and the detuning
//     delta = (freq + doppler - (Es_e - Es_g)) * 2 pi        (:63, :70-71)
// whose terms are ~4.3e14 Hz while delta is ~1e6-1e9 Hz; any reordering there
// moves delta by ~0.05 Hz and the Lorentzian by ~1e-8 relative.  Both are
// evaluated here with the reference's operation order and IEEE division
// (this file must be compiled with -ffp-contract=off; fused operations are
// written out explicitly where they are exact or harmless).
//
// Everything after delta is free to be restructured.  The population matrix
// of the reference (:86-97) is M = BR^T R^T - diag(sum_e R[g,e]), i.e.
// diagonal plus rank n_e, and its steady state p (M p = 0, sum p = 1) is the
// stationary law of a continuous-time Markov chain.  Instead of dense
// Gaussian elimination on the n_g x n_g system (:99-126, ~2 n_g^3/3 flop),
// the chain is reduced onto the n_e excited states:
//     K[e][f] = sum_g (S[g,e] / S_g) * BR[f][g]      (prob. excited f -> g -> e)
//     y = stationary vector of K (GTH state reduction, subtraction-free)
//     u_g = sum_f y_f BR[f][g];  p_g = (u_g / D_g) / sum_h (u_h / D_h)
// with S[g,e] = sum_beams s0 |d|^2 L, S_g = sum_e S[g,e] (the reference's
// s_total), D_g = Gamma/2 * Gamma_eff * S_g / (1 + S_g) the total excitation
// rate out of g, so F = sum_g p_g Fg_g and R = sum_g p_g D_g.
// When a ground state is dark (S_g = 0) or the excited chain is reducible
// the reduction does not apply; the point is then recomputed by
// rate_eval_dense, which restates the reference's pivot-skip elimination.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif
#include "mm_plan.h"

// 2.0 * numpy.pi, the reference's constant (rate_equations_jit.py:35,63,71)
#define MM_TWO_PI (2.0 * 3.141592653589793)
// RN(1 / MM_TWO_PI): the host compiler evaluates the IEEE division.
#define MM_INV_TWO_PI (1.0 / MM_TWO_PI)

// x / y correctly rounded from y_inv = RN(1/y) with one FMA correction
// (Markstein): q0 = RN(x y_inv), r = x - q0 y exactly, RN(q0 + r y_inv) is
// RN(x / y) for results clear of overflow and the subnormal range.  Three
// VALU operations in place of the scaled reciprocal-iteration sequence of
// an IEEE division, with the same result.
__device__ __forceinline__ double mm_div_rn(double x, double y, double y_inv) {
    const double q0 = x * y_inv;
    const double r = fma(-q0, y, x);
    return fma(r, y_inv, q0);
}

// 1/x for |x| far from the exponent limits (either sign).  v_rcp_f64 is good to
// ~4.6e-8 relative on gfx950; one Newton step brings it to ~2.2e-15, two to
// the correctly rounded value on every one of 4M sampled inputs in
// [1e10, 1e20].  Used only where the value passes through no cancellation
// afterwards.
__device__ __forceinline__ double mm_rcp1(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

__device__ __forceinline__ double mm_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    return r;
}

// Exact restatement of the reference evaluation (JIT flavour) for the
// corner cases the reduced solve does not cover.  Private arrays live in
// scratch; this path is not taken for any point whose ground states are all
// optically coupled.
template <int NG, int NE>
__device__ __attribute__((noinline)) void rate_eval_dense(const MMRatePlanDev* __restrict__ P,
                                                          double v, double z, double Bgrad,
                                                          double* F_out, double* R_out, double* p_out) {
    constexpr int NS = NG + NE;
    const double TWO_PI = MM_TWO_PI;
    double Es[NS];
    const double B_gauss = Bgrad * z * 1e4;
#pragma unroll 1
    for (int i = 0; i < NS; i++) Es[i] = P->E[i] + B_gauss * P->zd[i] * P->Gamma / TWO_PI;
    const double gamma_half = P->Gamma / 2.0;
    const double gamma_half_sq = gamma_half * gamma_half;
    double R_sum[NG][NE], F_tot[NG][NE], s_total[NG];
#pragma unroll 1
    for (int ig = 0; ig < NG; ig++) {
        s_total[ig] = 0.0;
#pragma unroll 1
        for (int ie = 0; ie < NE; ie++) R_sum[ig][ie] = F_tot[ig][ie] = 0.0;
    }
#pragma unroll 1
    for (int ib = 0; ib < P->n_b; ib++) {
        const double kdir = P->bkdir[ib], freq = P->bfreq[ib], s0 = P->bs0[ib];
        const int pol = P->bpol[ib];
        const double doppler = -P->k * kdir * v / TWO_PI;
#pragma unroll 1
        for (int ig = 0; ig < NG; ig++) {
#pragma unroll 1
            for (int ie = 0; ie < NE; ie++) {
                const double d2 = P->d2[ig][ie][pol];
                if (d2 < 1e-15) continue;
                const double omega_trans = Es[NG + ie] - Es[ig];
                const double delta_eff = (freq + doppler - omega_trans) * TWO_PI;
                const double L = gamma_half_sq / (delta_eff * delta_eff + gamma_half_sq);
                const double rate_unsat = gamma_half * s0 * d2 * L * P->Geff;
                R_sum[ig][ie] += rate_unsat;
                F_tot[ig][ie] += P->hbar * P->k * kdir * rate_unsat;
                s_total[ig] += s0 * d2 * L;
            }
        }
    }
#pragma unroll 1
    for (int ig = 0; ig < NG; ig++) {
        const double sat_corr = 1.0 / (1.0 + s_total[ig]);
#pragma unroll 1
        for (int ie = 0; ie < NE; ie++) {
            R_sum[ig][ie] *= sat_corr;
            F_tot[ig][ie] *= sat_corr;
        }
    }
    double A[NG][NG], b[NG], p[NG];
#pragma unroll 1
    for (int ig = 0; ig < NG; ig++) {
#pragma unroll 1
        for (int ik = 0; ik < NG; ik++) A[ig][ik] = 0.0;
    }
#pragma unroll 1
    for (int ig = 0; ig < NG; ig++) {
#pragma unroll 1
        for (int ie = 0; ie < NE; ie++) A[ig][ig] -= R_sum[ig][ie];
#pragma unroll 1
        for (int ik = 0; ik < NG; ik++)
#pragma unroll 1
            for (int ie = 0; ie < NE; ie++) A[ig][ik] += R_sum[ik][ie] * P->BR[ie][ig];
    }
#pragma unroll 1
    for (int j = 0; j < NG; j++) A[NG - 1][j] = 1.0;
#pragma unroll 1
    for (int i = 0; i < NG; i++) b[i] = 0.0;
    b[NG - 1] = 1.0;
#pragma unroll 1
    for (int col = 0; col < NG; col++) {
        double max_val = fabs(A[col][col]);
        int max_row = col;
#pragma unroll 1
        for (int row = col + 1; row < NG; row++) {
            if (fabs(A[row][col]) > max_val) {
                max_val = fabs(A[row][col]);
                max_row = row;
            }
        }
        if (max_row != col) {
#pragma unroll 1
            for (int j = 0; j < NG; j++) {
                const double t = A[col][j];
                A[col][j] = A[max_row][j];
                A[max_row][j] = t;
            }
            const double t = b[col];
            b[col] = b[max_row];
            b[max_row] = t;
        }
#pragma unroll 1
        for (int row = col + 1; row < NG; row++) {
            if (fabs(A[col][col]) < 1e-30) continue;
            const double factor = A[row][col] / A[col][col];
#pragma unroll 1
            for (int j = col; j < NG; j++) A[row][j] -= factor * A[col][j];
            b[row] -= factor * b[col];
        }
    }
#pragma unroll 1
    for (int i = 0; i < NG; i++) p[i] = 0.0;
#pragma unroll 1
    for (int i = NG - 1; i >= 0; i--) {
        double s = b[i];
#pragma unroll 1
        for (int j = i + 1; j < NG; j++) s -= A[i][j] * p[j];
        if (fabs(A[i][i]) > 1e-30) p[i] = s / A[i][i];
    }
#pragma unroll 1
    for (int i = 0; i < NG; i++)
        if (p[i] < 0) p[i] = 0.0;
    double total_p = 0.0;
#pragma unroll 1
    for (int i = 0; i < NG; i++) total_p += p[i];
    if (total_p > 0)
#pragma unroll 1
        for (int i = 0; i < NG; i++) p[i] /= total_p;
    double force = 0.0, R_scatter = 0.0;
#pragma unroll 1
    for (int ig = 0; ig < NG; ig++) {
#pragma unroll 1
        for (int ie = 0; ie < NE; ie++) {
            force += p[ig] * F_tot[ig][ie];
            R_scatter += p[ig] * R_sum[ig][ie];
        }
    }
    *F_out = force;
    *R_out = R_scatter;
    if (p_out)
#pragma unroll 1
        for (int i = 0; i < NG; i++) p_out[i] = p[i];
}

// Zeeman-shifted energies, reference operation order; states with a zero
// Zeeman coefficient keep their bare energy exactly, as E + (+-0) does.
template <int NG, int NE>
__device__ __forceinline__ void mm_shifted_energies(const MMRatePlanDev* __restrict__ P, double Bgrad,
                                                    double z, double (&Eg)[NG], double (&Ee)[NE]) {
    const double Bg = Bgrad * z * 1e4;
#pragma unroll
    for (int i = 0; i < NG; i++) {
        const double zdi = P->zd[i];
        Eg[i] = (zdi == 0.0) ? P->E[i] : P->E[i] + mm_div_rn(Bg * zdi * P->Gamma, MM_TWO_PI, MM_INV_TWO_PI);
    }
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const double zdi = P->zd[NG + i];
        Ee[i] = (zdi == 0.0) ? P->E[NG + i] : P->E[NG + i] + mm_div_rn(Bg * zdi * P->Gamma, MM_TWO_PI, MM_INV_TWO_PI);
    }
}

// One coupling term: s0 |d|^2 L for the detuning (fd - om) in Hz.  With
// C2 = s0 |d|^2 (Gamma/2)^2 / (2 pi)^2 and gq = (Gamma / (4 pi))^2 this is
// C2 / ((fd - om)^2 + gq), the reference's L (rate_equations_jit.py:71-72)
// with the 2 pi factor folded into the constants: fd - om, the only
// cancelling operation, is rounded exactly as the reference rounds it.
template <bool UNIT>
__device__ __forceinline__ double mm_term_w(const MMTerm& T, const MMRatePlanDev* __restrict__ P, double om,
                                            double v, double dop) {
    // freq + doppler: with kdir = +-1, kdir*dop is exact, so the fused form
    // rounds once exactly like the reference's sum
    const double fd = UNIT ? fma(T.kdir, dop, T.freq) : T.freq + mm_div_rn(P->mk * T.kdir * v, MM_TWO_PI, MM_INV_TWO_PI);
    const double d = fd - om;
    return T.C2 * mm_rcp1(fma(d, d, P->gq));
}

// Sum of s0 |d|^2 L (s) and of kdir s0 |d|^2 L (t) over the beams driving
// one (g, e) pair.  `dop` is ((-k) v)/(2 pi) when every |kdir| = 1.  Four
// terms per iteration: their scalar loads issue together and their
// dependency chains interleave.
template <bool UNIT>
__device__ __forceinline__ void mm_pair_sums(const MMRatePlanDev* __restrict__ P, int pair, double om,
                                             double v, double dop, double& s, double& t) {
    s = 0.0;
    t = 0.0;
    int it = P->pair_start[pair];
    const int t1 = P->pair_start[pair + 1];
    if (UNIT && P->pair_kind[pair] == 1) {
        // two +z then two -z beams: fd = freq +- dop rounds exactly as
        // fma(kdir, dop, freq) does, and one reciprocal serves all four
        // terms (simultaneous inversion: 1/x_a = x_b / (x_a x_b), ...)
        const MMTerm& A = P->terms[it];
        const MMTerm& B = P->terms[it + 1];
        const MMTerm& C = P->terms[it + 2];
        const MMTerm& D = P->terms[it + 3];
        const double da = (A.freq + dop) - om, db = (B.freq + dop) - om;
        const double dc = (C.freq - dop) - om, dd = (D.freq - dop) - om;
        const double xa = fma(da, da, P->gq), xb = fma(db, db, P->gq);
        const double xc = fma(dc, dc, P->gq), xd = fma(dd, dd, P->gq);
        const double pab = xa * xb, pcd = xc * xd;
        const double r = mm_rcp1(pab * pcd);
        const double rab = r * pcd, rcd = r * pab;
        const double wa = A.C2 * (xb * rab), wb = B.C2 * (xa * rab);
        const double wc = C.C2 * (xd * rcd), wd = D.C2 * (xc * rcd);
        const double sp = wa + wb, sm = wc + wd;
        s = sp + sm;
        t = sp - sm;
        return;
    }
    for (; it + 4 <= t1; it += 4) {
        const MMTerm& A = P->terms[it];
        const MMTerm& B = P->terms[it + 1];
        const MMTerm& C = P->terms[it + 2];
        const MMTerm& D = P->terms[it + 3];
        const double wa = mm_term_w<UNIT>(A, P, om, v, dop);
        const double wb = mm_term_w<UNIT>(B, P, om, v, dop);
        const double wc = mm_term_w<UNIT>(C, P, om, v, dop);
        const double wd = mm_term_w<UNIT>(D, P, om, v, dop);
        s += (wa + wb) + (wc + wd);
        t += fma(A.kdir, wa, B.kdir * wb) + fma(C.kdir, wc, D.kdir * wd);
    }
    for (; it < t1; ++it) {
        const MMTerm& A = P->terms[it];
        const double wa = mm_term_w<UNIT>(A, P, om, v, dop);
        s += wa;
        t = fma(A.kdir, wa, t);
    }
}

// Four-term pair from a packed record: two +z then two -z beams, one
// shared reciprocal (see mm_pair_sums).
__device__ __forceinline__ void mm_pair4(const MMPairRec& Q, double gq, double om, double dop, double& s,
                                         double& t) {
    const double da = (Q.f[0] + dop) - om, db = (Q.f[1] + dop) - om;
    const double dc = (Q.f[2] - dop) - om, dd = (Q.f[3] - dop) - om;
    const double xa = fma(da, da, gq), xb = fma(db, db, gq);
    const double xc = fma(dc, dc, gq), xd = fma(dd, dd, gq);
    // w_a + w_b = (c2_a x_b + c2_b x_a) / (x_a x_b), likewise for the -z pair;
    // both brought over the common denominator x_a x_b x_c x_d
    const double pab = xa * xb, pcd = xc * xd;
    const double r = mm_rcp1(pab * pcd);
    const double np = fma(Q.c2[0], xb, Q.c2[1] * xa) * pcd;
    const double nm = fma(Q.c2[2], xd, Q.c2[3] * xc) * pab;
    s = (np + nm) * r;
    t = (np - nm) * r;
}

// Stationary vector y (y[0] = 1, unnormalised) of the excited chain K
// (K[e][f]: probability excited f -> ground -> e) by GTH state reduction on
// the row-stochastic Q = K^T (Q[i][j] = K[j][i]): only off-diagonal entries
// are read, so no subtraction occurs.  K is overwritten; ok becomes false
// when the chain is reducible (a zero pivot sum), y is then meaningless.
template <int NE>
__device__ __forceinline__ void mm_gth_stationary(double (&K)[NE][NE], bool& ok, double (&y)[NE]) {
#pragma unroll
    for (int m = NE - 1; m >= 1; m--) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < m; j++) s += K[j][m];  // Q[m][j]
        ok = ok && (s > 0.0);
        const double rs = mm_rcp(s);
#pragma unroll
        for (int i = 0; i < m; i++) K[m][i] *= rs;  // Q[i][m] /= s
#pragma unroll
        for (int i = 0; i < m; i++)
#pragma unroll
            for (int j = 0; j < m; j++)
                if (i != j) K[j][i] = fma(K[m][i], K[j][m], K[j][i]);  // Q[i][j] += Q[i][m] Q[m][j]
    }
    y[0] = 1.0;
#pragma unroll
    for (int j = 1; j < NE; j++) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < j; i++) acc = fma(y[i], K[j][i], acc);  // Q[i][j]
        y[j] = acc;
    }
}

// Fast path.  Returns false (outputs untouched) when the reduction does not
// apply; the caller then uses rate_eval_dense.  `slab` is this thread's
// private LDS scratch of 2*NG doubles at stride `ls` (the block size), used
// for the per-ground rates so that the ground loop stays rolled and the
// register footprint small enough for several waves per SIMD.  `pop`, when
// not null, receives the NG populations (global memory, contiguous).
template <int NG, int NE, bool UNIT, int GU = 1, bool PK = false>
__device__ __forceinline__ bool rate_eval_fast(const MMRatePlanDev* __restrict__ P, double v, double z,
                                               double Bgrad, double& F, double& R, double* __restrict__ pop,
                                               double* slab, int ls) {
    double Ee[NE];
    const double Bg = Bgrad * z * 1e4;
#pragma unroll
    for (int i = 0; i < NE; i++) {
        const double zdi = P->zd[NG + i];
        Ee[i] = (zdi == 0.0) ? P->E[NG + i] : P->E[NG + i] + mm_div_rn(Bg * zdi * P->Gamma, MM_TWO_PI, MM_INV_TWO_PI);
    }
    const double dop = UNIT ? mm_div_rn(P->mk * v, MM_TWO_PI, MM_INV_TWO_PI) : 0.0;

    double K[NE][NE];  // K[e][f]: probability excited f -> (decay, re-excitation) -> e
#pragma unroll
    for (int e = 0; e < NE; e++)
#pragma unroll
        for (int f = 0; f < NE; f++) K[e][f] = 0.0;
    bool ok = true;

    // PK: the caller launched the variant for packed plans, so the general
    // pair path is not compiled into the ground-state loops at all
    constexpr bool PKC = PK && UNIT && NE <= 4;
    const bool packed = PKC || (UNIT && NE <= 4 && P->packed);
    const double gq = P->gq;
#pragma unroll GU
    for (int g = 0; g < NG; g++) {
        double S[NE];
        double Sg = 0.0, Tg = 0.0;
        double br[NE];
        if (packed) {
            // the ground record carries the pair mask and Gamma, so one
            // scalar load serves the shifted energy and the pair tests
            const MMGroundRec& G = P->grec[g];
            const MMPairRec* Q = &P->prec[g * 4];
            const double Eg = (G.zd == 0.0) ? G.E : G.E + mm_div_rn(Bg * G.zd * G.Gamma, MM_TWO_PI, MM_INV_TWO_PI);
            const unsigned mask = G.mask;
#pragma unroll
            for (int e = 0; e < NE; e++) {
                double s = 0.0, t = 0.0;
                if (mask & (1u << e)) mm_pair4(Q[e & 3], gq, Ee[e] - Eg, dop, s, t);
                S[e] = s;
                Sg += s;
                Tg += t;
                br[e] = G.br[e & 3];
            }
        } else {
            const double zdg = P->zd[g];
            const double Eg = (zdg == 0.0) ? P->E[g] : P->E[g] + mm_div_rn(Bg * zdg * P->Gamma, MM_TWO_PI, MM_INV_TWO_PI);
#pragma unroll
            for (int e = 0; e < NE; e++) {
                double s, t;
                mm_pair_sums<UNIT>(P, g * NE + e, Ee[e] - Eg, v, dop, s, t);
                S[e] = s;
                Sg += s;
                Tg += t;
                br[e] = P->BR[e][g];
            }
        }
        ok = ok && (Sg > 0.0);
        // 1/Sg and 1/(1+Sg) from one reciprocal
        const double s1 = 1.0 + Sg;
        const double rr = mm_rcp(Sg * s1);
        const double rS = rr * s1, sat = rr * Sg;
#pragma unroll
        for (int e = 0; e < NE; e++) {
            const double a = S[e] * rS;
#pragma unroll
            for (int f = 0; f < NE; f++)
                if (f != e) K[e][f] = fma(a, br[f], K[e][f]);
        }
        // 1 / D_g with D_g = Gamma/2 Geff Sg / (1 + Sg)
        slab[g * ls] = (rS * s1) * P->inv_gh_geff;
        slab[(NG + g) * ls] = P->fk_gh_geff * sat * Tg;   // Fg_g
    }

    double y[NE];
    mm_gth_stationary<NE>(K, ok, y);
    if (!ok) return false;

    double Z = 0.0, Fs = 0.0, Rs = 0.0;
#pragma unroll GU
    for (int g = 0; g < NG; g++) {
        double u = 0.0;
        if (packed) {
            const MMGroundRec& G = P->grec[g];
#pragma unroll
            for (int f = 0; f < NE; f++) u = fma(y[f], G.br[f & 3], u);
        } else {
#pragma unroll
            for (int f = 0; f < NE; f++) u = fma(y[f], P->BR[f][g], u);
        }
        const double q = u * slab[g * ls];
        Z += q;
        Fs = fma(q, slab[(NG + g) * ls], Fs);
        Rs += u;
        if (pop) slab[g * ls] = q;
    }
    const double rZ = mm_rcp(Z);
    F = Fs * rZ;
    R = Rs * rZ;
    if (pop) {
#pragma unroll 1
        for (int g = 0; g < NG; g++) pop[g] = slab[g * ls] * rZ;
    }
    return true;
}

// Evaluate one point: fast path, dense fallback for corner cases.  GU is the
// unroll factor of the fast path's ground-state loops (1 keeps the register
// footprint of the map kernels small); PK selects the variant for packed
// plans (the plan's `packed` flag must be set).
template <int NG, int NE, bool UNIT, int GU = 1, bool PK = false>
__device__ __forceinline__ void rate_eval(const MMRatePlanDev* __restrict__ P, double v, double z, double Bgrad,
                                          double& F, double& R, double* __restrict__ pop, double* slab, int ls) {
    bool ok = false;
    if (!P->force_dense) ok = rate_eval_fast<NG, NE, UNIT, GU, PK>(P, v, z, Bgrad, F, R, pop, slab, ls);
    if (!ok) {
        // through temporaries: taking the address of F and R would keep the
        // caller's force and rate in scratch memory on every call, fallback
        // or not
        double Fd, Rd;
        rate_eval_dense<NG, NE>(P, v, z, Bgrad, &Fd, &Rd, pop);
        F = Fd;
        R = Rd;
    }
}
