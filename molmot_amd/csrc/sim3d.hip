// 3D rate-equation force with the full ground-state Zeeman Hamiltonian
// (molmot/mot/simulator_3d.py:168-365) and the leapfrog ensemble
// (simulate_trajectory_3d / simulate_ensemble_3d, :439-532), on gfx950.
//
// Layout: one 16-lane group per point (or particle), four groups per
// 64-lane block (one wavefront).  Each group owns an LDS slab:
//   A [NG][NG] complex   Hermitised H_Z = Bx Zx + By Zy + Bz Zz, diagonalised in place
//   V [NG][NG] complex   the accumulated rotations (eigenvectors in the columns);
//                        U^+ |d| = V^H |d| is formed after convergence.  Along a
//                        trajectory the next step starts from A = V^H H V.
// Diagonalisation: cyclic complex Jacobi with the round-robin ("tournament")
// ordering, NG/2 disjoint rotations per round applied in parallel by the
// group's lanes; sweeps until the squared off-diagonal norm falls below
// 1e-24 of the squared Frobenius norm (at most MAXSWEEP; the eigenvectors are
// then good to ~1e-12 of the gap, the eigenvalues to ~1e-24).  Eigenvalues are then ranked ascending
// (numpy's eigh order) and eigenvalue i is paired with bare energy i, as the
// reference does.  The rate equations follow the reference's two passes
// (saturation sum, then rates), the 12 x 12 generator with its last row
// replaced by ones is solved by row-parallel Gaussian elimination with partial
// pivoting, populations are clamped and normalised.  FP64 throughout.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "mm_errors.h"

#define S3_MAXG 12
#define S3_MAXE 4
#define S3_MAXB 48
#define S3_GROUP 16
#define S3_PTS 4
#define S3_MAXSWEEP 12

struct S3Plan {
    int ng, ne, nb, pad;
    double Gamma, k, Geff, hbar, mass;
    double E[S3_MAXG + S3_MAXE];
    double zr[3][S3_MAXG][S3_MAXG], zi[3][S3_MAXG][S3_MAXG];  // ground Zeeman blocks
    double zde[S3_MAXE];                                      // Re diag of the excited Zz block
    double dabs[S3_MAXG][3 * S3_MAXE];                        // |tdm[g, ng+e, q]|, column e*3+q
    double bk[S3_MAXB][3], bfreq[S3_MAXB], bs0[S3_MAXB], bw[S3_MAXB];
    int bq[S3_MAXB];
    int pairs[S3_MAXG][S3_MAXG / 2 + 1][2];  // round-robin schedule: [round][slot] -> (p, q), -1 = none
    int n_rounds, n_slots;
    // spin-rotation eigenbasis (mm_force3d_plan_set_spin_basis); spin = 0: Jacobi
    int spin;
    int cls[4];                   // beams of lab polarisation q: perm[cls[q] .. cls[q + 1])
    int perm[S3_MAXB];
    double pk[S3_MAXB][3], pf[S3_MAXB], ps0[S3_MAXB], pwr[S3_MAXB];  // beam tables in perm order
    int nd;                       // distinct (direction, radius) pairs of the beams
    int pdir[S3_MAXB];            // pair of permuted beam t
    double dk[S3_MAXB][3], dw[S3_MAXB];
    double lam0[S3_MAXG];         // eigenvalues of the ground Zz block, ascending
    double sd[S3_MAXG][3 * S3_MAXE][6];  // per (eigenvector g, column c): W^H|d|, W^H 2Sx|d|, W^H 2Sy|d| (re, im)
    double inv_tot[S3_MAXE];      // 1 / sum_{g,q} |d[g, e, q]|^2 (0 where that sum is <= 1e-30)
};

struct mm_force3d_plan {
    S3Plan host;
    S3Plan* dev;
};

template <int NG, int NE>
struct S3Post {  // rate-equation data, formed once A is diagonal
    double BR[NE][NG];
    double Es_g[NG], Es_e[NE];
    double stot[NG], Rt[NG][NE], FK[NG][3], p[NG];
};

template <int NG, int NE>
struct S3Slab {
    union {             // A until its eigenvalues are ranked, then the rate data
        double2 A[NG][NG];
        S3Post<NG, NE> q;
    } a;
    double2 V[NG][NG];  // accumulated rotations: eigenvectors in the columns
    double prm[NG / 2 + 1][6];  // c, s, s e, c e per slot
    double lam[NG];
    double pw[3][3];
    double red[S3_GROUP];
    int flag;
    union {                   // live in turn: dsq (rates), M (solve)
        double dsq[NG][3 * NE];
        double M[NG][NG + 1];
    } u;
};

// (x, y) <- (c x - w y, s x + u y) for complex w = s e, u = c e: one
// rotation of a pair of entries, written as fused multiply-adds (the library
// is compiled with contraction off).
__device__ __forceinline__ void s3_rot(double2& x, double2& y, double c, double s, double2 w, double2 u) {
    const double2 a = x, b = y;
    x = make_double2(fma(c, a.x, fma(-w.x, b.x, w.y * b.y)), fma(c, a.y, fma(-w.x, b.y, -(w.y * b.x))));
    y = make_double2(fma(s, a.x, fma(u.x, b.x, -(u.y * b.y))), fma(s, a.y, fma(u.x, b.y, u.y * b.x)));
}

__device__ __forceinline__ double group_sum(double x) {
#pragma unroll
    for (int o = 8; o >= 1; o >>= 1) x += __shfl_xor(x, o, S3_GROUP);
    return x;
}

// 1/sqrt(x) and 1/x for x > 0 away from the exponent limits: hardware
// estimate and two Newton steps (relative error ~1e-16).
__device__ __forceinline__ double s3_rsqrt(double x) {
    double y = __builtin_amdgcn_rsq(x);
    const double hx = 0.5 * x;
    y = y * fma(-hx * y, y, 1.5);
    y = y * fma(-hx * y, y, 1.5);
    return y;
}
__device__ __forceinline__ double s3_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

// Round-robin ("tournament") schedule over m = NG rounded up to even players,
// player m-1 fixed: in round rd, slot 0 pairs (m-1, rd) and slot sl pairs
// (rd + sl, rd - sl) mod (m-1).  Pairs that involve the padding player are
// (-1, -1).  Computed in registers: the schedule is the same for every
// point and needs no memory.
template <int NG>
__device__ __forceinline__ void s3_pair(int rd, int sl, int& p, int& q) {
    constexpr int m = (NG + 1) / 2 * 2;
    int a, b;
    if (sl == 0) {
        a = m - 1;
        b = rd;
    } else {
        a = rd + sl;
        a -= (a >= m - 1) ? (m - 1) : 0;
        b = rd - sl;
        b += (b < 0) ? (m - 1) : 0;
    }
    if (a >= NG || b >= NG) a = b = -1;
    p = a < b ? a : b;
    q = a < b ? b : a;
}

// numpy's sum order for n <= 15 elements (8 accumulators, then the tail)
template <int N>
__device__ __forceinline__ double np_sum(const double* a) {
    if constexpr (N < 8) {
        double s = a[0];
        for (int i = 1; i < N; i++) s += a[i];
        return s;
    } else {
        double r = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
        for (int i = 8; i < N; i++) r += a[i];
        return r;
    }
}

// Force at one point; called by all 16 lanes of a group, with block-wide
// barriers (the block is one wavefront).  Returns F, R on every lane and
// leaves p in the slab.
// warm: start the diagonalisation from the eigenvectors V left by the
// previous call on this slab (A <- V^H H V), which along a trajectory is
// close to diagonal already; otherwise from V = I.
template <int NG, int NE>
__device__ void s3_force(const S3Plan* __restrict__ P, S3Slab<NG, NE>& S, int lane, const double (&r)[3],
                         const double (&v)[3], double Bg, double (&F)[3], double& Rsc, bool warm) {
    constexpr int NC = 3 * NE;
    // 1. quadrupole field (simulator_3d.py:209-219)
    const double Bx = -Bg * r[0] / 2.0 * 1e4, By = -Bg * r[1] / 2.0 * 1e4, Bz = Bg * r[2] * 1e4;
    const double Bm = sqrt(Bx * Bx + By * By + Bz * Bz);
    double bhx = 0.0, bhy = 0.0, bhz = 1.0;
    if (Bm > 1e-10) {
        bhx = Bx / Bm;
        bhy = By / Bm;
        bhz = Bz / Bm;
    }
    // 2. Hermitised ground Zeeman block and the |d| columns
    for (int t = lane; t < NG * NG; t += S3_GROUP) {
        const int i = t / NG, j = t % NG;
        const double hr = (Bx * P->zr[0][i][j] + By * P->zr[1][i][j]) + Bz * P->zr[2][i][j];
        const double hi = (Bx * P->zi[0][i][j] + By * P->zi[1][i][j]) + Bz * P->zi[2][i][j];
        const double gr = (Bx * P->zr[0][j][i] + By * P->zr[1][j][i]) + Bz * P->zr[2][j][i];
        const double gi = (Bx * P->zi[0][j][i] + By * P->zi[1][j][i]) + Bz * P->zi[2][j][i];
        S.a.A[i][j] = make_double2(0.5 * (hr + gr), 0.5 * (hi - gi));
    }
    if (!warm)
        for (int t = lane; t < NG * NG; t += S3_GROUP)
            S.V[t / NG][t % NG] = make_double2((t / NG == t % NG) ? 1.0 : 0.0, 0.0);
    if (lane < 3) {
        // squared Wigner d^1 column q_lab = lane, normalised (simulator_3d.py:111-165)
        double w0, w1, w2;
        const double nb = sqrt(bhx * bhx + bhy * bhy + bhz * bhz);
        if (nb < 1e-15) {
            w0 = w1 = w2 = 1.0 / 3.0;
        } else {
            const double ct = fmin(fmax(bhz / nb, -1.0), 1.0);
            const double st = sin(acos(ct));
            const double a = (1 + ct) / 2, b = (1 - ct) / 2, c = st / sqrt(2.0);
            const double m[3][3] = {{a * a, c * c, b * b}, {c * c, ct * ct, c * c}, {b * b, c * c, a * a}};
            w0 = m[0][lane];
            w1 = m[1][lane];
            w2 = m[2][lane];
            const double s = (w0 + w1) + w2;
            if (s > 1e-15) {
                w0 /= s;
                w1 /= s;
                w2 /= s;
            }
        }
        S.pw[lane][0] = w0;
        S.pw[lane][1] = w1;
        S.pw[lane][2] = w2;
    }
    __syncthreads();
    if (warm) {
        // T = H V, then A = V^H T; lane j forms column j of both, and keeps
        // its column of T in registers
        double2 tj[NG];
        if (lane < NG) {
            const int j = lane;
            double2 vj[NG];
#pragma unroll
            for (int m = 0; m < NG; m++) vj[m] = S.V[m][j];
#pragma unroll
            for (int k = 0; k < NG; k++) {
                double xr = 0.0, xi = 0.0;
#pragma unroll
                for (int m = 0; m < NG; m++) {
                    const double2 h = S.a.A[k][m];
                    xr = fma(h.x, vj[m].x, fma(-h.y, vj[m].y, xr));
                    xi = fma(h.x, vj[m].y, fma(h.y, vj[m].x, xi));
                }
                tj[k] = make_double2(xr, xi);
            }
        }
        __syncthreads();  // H is read by every lane before A is overwritten
        if (lane < NG) {
            const int j = lane;
#pragma unroll 1
            for (int i = 0; i < NG; i++) {
                double xr = 0.0, xi = 0.0;
#pragma unroll
                for (int k = 0; k < NG; k++) {
                    const double2 w = S.V[k][i];  // conj(w) tj[k]
                    xr = fma(w.x, tj[k].x, fma(w.y, tj[k].y, xr));
                    xi = fma(w.x, tj[k].y, fma(-w.y, tj[k].x, xi));
                }
                S.a.A[i][j] = make_double2(xr, xi);
            }
        }
        __syncthreads();
    }
    // 3. cyclic Jacobi, round-robin order
    for (int sweep = 0; sweep < S3_MAXSWEEP; sweep++) {
        double off = 0.0, fro = 0.0;
        for (int t = lane; t < NG * NG; t += S3_GROUP) {
            const int i = t / NG, j = t % NG;
            const double2 a = S.a.A[i][j];
            const double m2 = a.x * a.x + a.y * a.y;
            fro += m2;
            if (i != j) off += m2;
        }
        off = group_sum(off);
        fro = group_sum(fro);
        const int done = !(off > 1e-24 * fro) ? 1 : 0;
        if (__syncthreads_and(done)) break;
        constexpr int NR = NG > 1 ? (NG + 1) / 2 * 2 - 1 : 0, NS = (NG + 1) / 2;
        for (int rd = 0; rd < NR; rd++) {
            if (lane < NS) {
                int p, q;
                s3_pair<NG>(rd, lane, p, q);
                double c = 1.0, s = 0.0, er = 1.0, ei = 0.0;
                if (p >= 0) {
                    const double2 apq = S.a.A[p][q];
                    const double m2 = apq.x * apq.x + apq.y * apq.y;
                    if (m2 > 1e-300) {
                        // 1/|a_pq| by rsqrt with two Newton steps (~1e-16): the
                        // rotation needs no correctly rounded quotients
                        const double ia = s3_rsqrt(m2);
                        const double tau = (S.a.A[q][q].x - S.a.A[p][p].x) * (0.5 * ia);
                        const double den = fabs(tau) + sqrt(fma(tau, tau, 1.0));
                        const double tt = (tau >= 0.0 ? 1.0 : -1.0) * s3_rcp(den);
                        c = s3_rsqrt(fma(tt, tt, 1.0));
                        s = tt * c;
                        er = apq.x * ia;
                        ei = apq.y * ia;
                    }
                }
                S.prm[lane][0] = c;
                S.prm[lane][1] = s;
                S.prm[lane][2] = s * er;
                S.prm[lane][3] = s * ei;
                S.prm[lane][4] = c * er;
                S.prm[lane][5] = c * ei;
            }
            __syncthreads();
            // columns of A and V: [X_kp, X_kq] <- [c X_kp - s e^{-i phi} X_kq, s X_kp + c e^{-i phi} X_kq];
            // lane k takes row k through the round's disjoint rotations
            if (lane < NG) {
                const int k = lane;
#pragma unroll
                for (int sl = 0; sl < NS; sl++) {
                    int p, q;
                    s3_pair<NG>(rd, sl, p, q);
                    if (p < 0) continue;
                    const double c = S.prm[sl][0], s = S.prm[sl][1];
                    // e^{-i phi}: the conjugates of s e and c e
                    const double2 w = make_double2(S.prm[sl][2], -S.prm[sl][3]);
                    const double2 u = make_double2(S.prm[sl][4], -S.prm[sl][5]);
                    s3_rot(S.a.A[k][p], S.a.A[k][q], c, s, w, u);
                    s3_rot(S.V[k][p], S.V[k][q], c, s, w, u);
                }
            }
            __syncthreads();
            // rows of A: [A_pk, A_qk] <- [c A_pk - s e^{i phi} A_qk, s A_pk + c e^{i phi} A_qk];
            // lane k takes column k
            if (lane < NG) {
                const int kk = lane;
#pragma unroll
                for (int sl = 0; sl < NS; sl++) {
                    int p, q;
                    s3_pair<NG>(rd, sl, p, q);
                    if (p < 0) continue;
                    const double c = S.prm[sl][0], s = S.prm[sl][1];
                    const double2 w = make_double2(S.prm[sl][2], S.prm[sl][3]);
                    const double2 u = make_double2(S.prm[sl][4], S.prm[sl][5]);
                    s3_rot(S.a.A[p][kk], S.a.A[q][kk], c, s, w, u);
                }
            }
            __syncthreads();
        }
    }
    // 4. rank eigenvalues ascending; |U^+ d|^2 in eigenvalue order
    if (lane < NG) {
        const double li = S.a.A[lane][lane].x;
        int rk = 0;
        for (int j = 0; j < NG; j++) {
            const double lj = S.a.A[j][j].x;
            rk += (lj < li || (lj == li && j < lane)) ? 1 : 0;
        }
        S.lam[rk] = li;
        // (V^H |d|)[lane][c]: |d| is real
        double2 vc[NG];
#pragma unroll
        for (int g = 0; g < NG; g++) vc[g] = S.V[g][lane];
#pragma unroll 1
        for (int c = 0; c < NC; c++) {
            double xr = 0.0, xi = 0.0;
#pragma unroll
            for (int g = 0; g < NG; g++) {
                const double d = P->dabs[g][c];
                xr = fma(vc[g].x, d, xr);
                xi = fma(-vc[g].y, d, xi);
            }
            S.u.dsq[rk][c] = xr * xr + xi * xi;
        }
    }
    __syncthreads();
    const double G = P->Gamma;
    if (lane < NG) S.a.q.Es_g[lane] = P->E[lane] + S.lam[lane] * G / (2 * M_PI);
    if (lane < NE) {
        S.a.q.Es_e[lane] = P->E[NG + lane] + Bz * P->zde[lane] * G / (2 * M_PI);
        // branching ratios in the eigenbasis (simulator_3d.py:252-259)
        double tot = 0.0;
        for (int g = 0; g < NG; g++)
            for (int q = 0; q < 3; q++) tot += S.u.dsq[g][lane * 3 + q];
        for (int g = 0; g < NG; g++) {
            double sg = 0.0;
            for (int q = 0; q < 3; q++) sg += S.u.dsq[g][lane * 3 + q];
            S.a.q.BR[lane][g] = tot > 1e-30 ? sg / tot : 0.0;
        }
    }
    __syncthreads();
    // 5. rates, lane = ground state.  The reference's two passes over the
    // beams (saturation sum, then rates divided by 1 + s_total) are one pass
    // here: every rate of ground state g carries the same factor
    // Gamma/2 Geff / (1 + s_total_g), applied once after the sums.  A beam's
    // detuning does not depend on q, so each (beam, e) needs one Lorentzian;
    // |d|^2 below 1e-15 is dropped, as the reference skips it.
    const double gh = G / 2.0, gh2 = gh * gh;
    const double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    if (lane < NG) {
        const int g = lane;
        double om[NE], d2[NE][3];
#pragma unroll
        for (int e = 0; e < NE; e++) {
            om[e] = S.a.q.Es_e[e] - S.a.q.Es_g[g];
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const double d = S.u.dsq[g][e * 3 + q];
                d2[e][q] = d < 1e-15 ? 0.0 : d;
            }
        }
        double st = 0.0, Rt[NE], fk[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int e = 0; e < NE; e++) Rt[e] = 0.0;
        for (int b = 0; b < P->nb; b++) {
            const double kx = P->bk[b][0], ky = P->bk[b][1], kz = P->bk[b][2];
            const double dop = -P->k * (kx * v[0] + ky * v[1] + kz * v[2]) / (2 * M_PI);
            double gauss = 1.0;
            if (P->bw[b] > 0.0) {
                const double rpar = r[0] * kx + r[1] * ky + r[2] * kz;
                const double rp2 = fmax(rr - rpar * rpar, 0.0);
                gauss = exp(-2.0 * rp2 / (P->bw[b] * P->bw[b]));
            }
            const double se = P->bs0[b] * gauss;
            const int ql = P->bq[b];
            const double w0 = S.pw[ql][0], w1 = S.pw[ql][1], w2 = S.pw[ql][2];
            const double fb = P->bfreq[b] + dop;
            double sb = 0.0;
#pragma unroll
            for (int e = 0; e < NE; e++) {
                const double de = (fb - om[e]) * (2 * M_PI);
                const double x = de * de + gh2;
                double L = __builtin_amdgcn_rcp(x);
                L = fma(L, fma(-x, L, 1.0), L);
                L *= gh2;
                const double rs = L * ((w0 * d2[e][0] + w1 * d2[e][1]) + w2 * d2[e][2]);
                sb = fma(se, rs, sb);
                Rt[e] = fma(rs, se, Rt[e]);
            }
            st += sb;
            fk[0] = fma(kx, sb, fk[0]);
            fk[1] = fma(ky, sb, fk[1]);
            fk[2] = fma(kz, sb, fk[2]);
        }
        S.a.q.stot[g] = st;
        const double sc = gh * P->Geff / (1.0 + st);
#pragma unroll
        for (int e = 0; e < NE; e++) S.a.q.Rt[g][e] = sc * Rt[e];
        S.a.q.FK[g][0] = sc * fk[0];
        S.a.q.FK[g][1] = sc * fk[1];
        S.a.q.FK[g][2] = sc * fk[2];
    }
    __syncthreads();
    // 6. generator with last row = 1 (simulator_3d.py:318-331); lane = row
    if (lane < NG) {
        const int g = lane;
        double sg = S.a.q.Rt[g][0];
        for (int e = 1; e < NE; e++) sg += S.a.q.Rt[g][e];
        for (int k2 = 0; k2 < NG; k2++) {
            double m = (k2 == g) ? -sg : 0.0;
            for (int e = 0; e < NE; e++) m += S.a.q.Rt[k2][e] * S.a.q.BR[e][g];
            S.u.M[g][k2] = (g == NG - 1) ? 1.0 : m;
        }
        S.u.M[g][NG] = (g == NG - 1) ? 1.0 : 0.0;
    }
    if (lane == 0) S.flag = 0;
    __syncthreads();
    // Gaussian elimination with partial pivoting (first maximum), rows in parallel
    for (int c = 0; c < NG; c++) {
        double best = -1.0;
        int piv = c;
        for (int i = c; i < NG; i++) {
            const double a = fabs(S.u.M[i][c]);
            if (a > best) {
                best = a;
                piv = i;
            }
        }
        if (best == 0.0 && lane == 0) S.flag = 1;
        __syncthreads();
        if (piv != c) {
            for (int j = lane; j <= NG; j += S3_GROUP) {
                const double t = S.u.M[c][j];
                S.u.M[c][j] = S.u.M[piv][j];
                S.u.M[piv][j] = t;
            }
        }
        __syncthreads();
        if (lane > c && lane < NG && S.u.M[c][c] != 0.0) {
            const double l = S.u.M[lane][c] / S.u.M[c][c];
            for (int j = c + 1; j <= NG; j++) S.u.M[lane][j] -= l * S.u.M[c][j];
            S.u.M[lane][c] = 0.0;
        }
        __syncthreads();
    }
    if (lane == 0) {
        double x[NG];
        if (S.flag) {
            for (int i = 0; i < NG; i++) x[i] = 1.0 / NG;
        } else {
            for (int i = NG - 1; i >= 0; i--) {
                double acc = S.u.M[i][NG];
                for (int j = i + 1; j < NG; j++) acc -= S.u.M[i][j] * x[j];
                x[i] = acc / S.u.M[i][i];
            }
            for (int i = 0; i < NG; i++) x[i] = fmax(x[i], 0.0);
            const double ps = np_sum<NG>(x);
            for (int i = 0; i < NG; i++) x[i] = ps > 0.0 ? x[i] / ps : 1.0 / NG;
        }
        for (int i = 0; i < NG; i++) S.a.q.p[i] = x[i];
    }
    __syncthreads();
    // 7. force and scattering rate
    double fx = 0.0, fy = 0.0, fz = 0.0, rs = 0.0;
    if (lane < NG) {
        const double pg = S.a.q.p[lane];
        fx = S.a.q.FK[lane][0] * pg;
        fy = S.a.q.FK[lane][1] * pg;
        fz = S.a.q.FK[lane][2] * pg;
        double sg = 0.0;
        for (int e = 0; e < NE; e++) sg += S.a.q.Rt[lane][e];
        rs = sg * pg;
    }
    const double hk = P->hbar * P->k;
    F[0] = hk * group_sum(fx);
    F[1] = hk * group_sum(fy);
    F[2] = hk * group_sum(fz);
    Rsc = group_sum(rs);
}

template <int NG, int NE>
__global__ __launch_bounds__(64) void k_force3d(const S3Plan* __restrict__ P, long n, const double* __restrict__ r_in,
                                                const double* __restrict__ v_in, double Bg, double* __restrict__ F_out,
                                                double* __restrict__ R_out, double* __restrict__ pop_out) {
    __shared__ S3Slab<NG, NE> slabs[S3_PTS];
    const int grp = threadIdx.x / S3_GROUP, lane = threadIdx.x % S3_GROUP;
    const long i = (long)blockIdx.x * S3_PTS + grp;
    const long ii = i < n ? i : n - 1;
    double r[3], v[3], F[3], R;
    for (int a = 0; a < 3; a++) {
        r[a] = r_in[ii * 3 + a];
        v[a] = v_in[ii * 3 + a];
    }
    s3_force<NG, NE>(P, slabs[grp], lane, r, v, Bg, F, R, false);
    if (i < n) {
        if (lane < 3) F_out[i * 3 + lane] = F[lane];
        if (lane == 0 && R_out) R_out[i] = R;
        if (pop_out && lane < NG) pop_out[i * NG + lane] = slabs[grp].a.q.p[lane];
    }
}

// Leapfrog ensemble (simulator_3d.py:439-502): r_j = r_{j-1} + dt v_half,
// escape test, F(r_j, v_half), v_half += dt F / m, stored v = mean of the
// two half-step velocities; after an escape the last r and v_half fill the
// rest of the stored trajectory.
template <int NG, int NE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_leapfrog3d(const S3Plan* __restrict__ P, long n,
                                                   const double* __restrict__ r0, const double* __restrict__ v0,
                                                   double Bg, double dt, int n_steps, double r_escape,
                                                   double* __restrict__ r_traj, double* __restrict__ v_traj,
                                                   double* __restrict__ r_fin, double* __restrict__ v_fin,
                                                   int* __restrict__ n_act) {
    __shared__ S3Slab<NG, NE> slabs[S3_PTS];
    const int grp = threadIdx.x / S3_GROUP, lane = threadIdx.x % S3_GROUP;
    const long i = (long)blockIdx.x * S3_PTS + grp;
    const bool live = i < n;
    const long ii = live ? i : n - 1;
    const double m = P->mass;
    double r[3], vh[3], F[3], R;
    for (int a = 0; a < 3; a++) {
        r[a] = r0[ii * 3 + a];
        vh[a] = v0[ii * 3 + a];
    }
    if (live && r_traj && lane < 3) {
        r_traj[(i * n_steps) * 3 + lane] = r[lane];
        v_traj[(i * n_steps) * 3 + lane] = vh[lane];
    }
    s3_force<NG, NE>(P, slabs[grp], lane, r, vh, Bg, F, R, false);
    for (int a = 0; a < 3; a++) vh[a] = vh[a] + 0.5 * dt * F[a] / m;
    int esc = 0, nact = n_steps;
    for (int j = 1; j < n_steps; j++) {
        double rn[3];
        if (!esc) {
            for (int a = 0; a < 3; a++) rn[a] = r[a] + dt * vh[a];
            const double nr = sqrt(rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2]);
            for (int a = 0; a < 3; a++) r[a] = rn[a];
            if (nr > r_escape) {
                esc = 1;
                nact = j;
            }
        }
        // every group keeps calling the force so that the block's barriers line up
        double Fj[3], Rj;
        s3_force<NG, NE>(P, slabs[grp], lane, r, vh, Bg, Fj, Rj, true);
        double vs[3];
        if (!esc) {
            for (int a = 0; a < 3; a++) {
                const double vn = vh[a] + dt * Fj[a] / m;
                vs[a] = 0.5 * (vh[a] + vn);
                vh[a] = vn;
            }
        } else {
            for (int a = 0; a < 3; a++) vs[a] = vh[a];
        }
        if (live && r_traj && lane < 3) {
            r_traj[(i * n_steps + j) * 3 + lane] = r[lane];
            v_traj[(i * n_steps + j) * 3 + lane] = vs[lane];
        }
        if (__syncthreads_and(esc || !live)) {
            // the whole block has escaped: fill the remaining steps
            if (live && r_traj && lane < 3)
                for (int jj = j + 1; jj < n_steps; jj++) {
                    r_traj[(i * n_steps + jj) * 3 + lane] = r[lane];
                    v_traj[(i * n_steps + jj) * 3 + lane] = vh[lane];
                }
            break;
        }
    }
    if (live) {
        if (lane < 3) {
            if (r_fin) r_fin[i * 3 + lane] = r[lane];
            if (v_fin) v_fin[i * 3 + lane] = vh[lane];
        }
        if (lane == 0 && n_act) n_act[i] = nact;
    }
}

// ---------------------------------------------------------------------------
// Spin-rotation eigenbasis.  When the ground Zeeman blocks are s times the
// spin operators of a direct sum of spin-1/2 representations
// ([Zx, Zy] = i s Zz and cyclic, Zx^2 + Zy^2 + Zz^2 = 3 s^2 / 4; the SrOH
// tables, to 2e-12), H_Z = s |B| n.S with n = B / |B|, and its eigenvectors
// follow in closed form from those of Zz (W, eigenvalues lam0, ascending):
// U = R W with R = exp(-i theta m.S) the rotation taking z to n about
// m = z x n / |z x n|.  Since (2 m.S)^2 = 1,
//     U^H |d| = W^H R^H |d| = c W^H|d| + i (al W^H 2Sx|d| + be W^H 2Sy|d|),
// c = cos(theta/2) = sqrt((1 + n_z)/2), (al, be) = sin(theta/2) (m_x, m_y)
// = (-n_y, n_x) / (2 c); at n = -z, m = y (al = 0, be = 1).  The eigenvalues
// are |B| lam0 in the same (ascending) order, paired with bare energy i as
// the reference pairs them.  Wherever the spectrum is degenerate (SrOH: two
// six-fold eigenvalues) this fixes the basis of each eigenspace that the
// reference leaves to its eigensolver: a continuous function of n except at
// n = -z; at B = 0 exactly the basis is the identity, as eigh returns for
// the zero matrix.  The per-point work is then a linear combination of three
// constant tables, so one lane takes one point: no LDS matrices, no
// barriers.
//
// Steady state: with R~_ge the unscaled rates of ground state g into e
// (sum over beams of s_eff L |<e|d.eps|g>|^2 weights) and D_g its total
// out-rate, p_g = u_g / D_g / Z with u_g = sum_f y_f BR_fg, where y is the
// stationary vector of K_ef = sum_g (R~_ge / R~_g) BR_fg (the excited chain,
// solved by GTH state reduction).  F, R and Z are sums over g of u_g times
// per-g values, so they are accumulated per excited state f in the same
// pass over g and contracted with y at the end.  Branching ratios use
// sum_{g,q} |(U^H d)_{g,eq}|^2 = sum_{g,q} |d_{g,eq}|^2 (U unitary).  A
// point whose chain is reducible (a ground state no beam couples, or an
// excited state no ground state feeds) takes the generator solve of the
// reference instead (s3_spin_dense).
// ---------------------------------------------------------------------------
struct S3SpinFrame {
    double Bm, Bz, nx, ny, nz, c, al, be;
    bool ident;           // B = 0 exactly: the eigenvectors of the zero matrix are the identity (eigh's)
    double pw[3][3];      // pw[q_lab][q]: squared Wigner d^1 weights (simulator_3d.py:111-165)
    double Ese[S3_MAXE];  // Zeeman-shifted excited energies (Hz)
};

template <int NE>
__device__ __forceinline__ void s3_spin_frame(const S3Plan* __restrict__ P, const double (&r)[3], double Bg,
                                              S3SpinFrame& fr) {
    const double Bx = -Bg * r[0] / 2.0 * 1e4, By = -Bg * r[1] / 2.0 * 1e4, Bz = Bg * r[2] * 1e4;
    const double Bm = sqrt(Bx * Bx + By * By + Bz * Bz);
    double nx = 0.0, ny = 0.0, nz = 1.0;
    if (Bm > 1e-10) {
        nx = Bx / Bm;
        ny = By / Bm;
        nz = Bz / Bm;
    }
    fr.Bm = Bm;
    fr.Bz = Bz;
    fr.ident = !(Bm > 0.0);
    fr.nx = nx;
    fr.ny = ny;
    fr.nz = nz;
    const double c = sqrt(fmax(0.0, 0.5 * (1.0 + nz)));
    if (c > 1e-8) {
        const double h = 0.5 / c;
        fr.c = c;
        fr.al = -ny * h;
        fr.be = nx * h;
    } else {
        fr.c = 0.0;
        fr.al = 0.0;
        fr.be = 1.0;
    }
    const double nb = sqrt(nx * nx + ny * ny + nz * nz);
    if (nb < 1e-15) {
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) fr.pw[a][b] = 1.0 / 3.0;
    } else {
        const double ct = fmin(fmax(nz / nb, -1.0), 1.0);
        const double st = sin(acos(ct));
        const double a = (1 + ct) / 2, b = (1 - ct) / 2, cc = st / sqrt(2.0);
        const double m[3][3] = {{a * a, cc * cc, b * b}, {cc * cc, ct * ct, cc * cc}, {b * b, cc * cc, a * a}};
#pragma unroll
        for (int ql = 0; ql < 3; ql++) {
            double w0 = m[0][ql], w1 = m[1][ql], w2 = m[2][ql];
            const double sm = (w0 + w1) + w2;
            if (sm > 1e-15) {
                w0 /= sm;
                w1 /= sm;
                w2 /= sm;
            }
            fr.pw[ql][0] = w0;
            fr.pw[ql][1] = w1;
            fr.pw[ql][2] = w2;
        }
    }
    const double G2 = P->Gamma / (2 * M_PI);
#pragma unroll
    for (int e = 0; e < NE; e++) fr.Ese[e] = P->E[P->ng + e] + Bz * P->zde[e] * G2;
}

// |(U^H |d|)_{g, c}|^2 for the NC = 3 NE columns of eigenvector g
template <int NE>
__device__ __forceinline__ void s3_spin_dsq(const S3Plan* __restrict__ P, const S3SpinFrame& fr, int g,
                                            double (&dsq)[3 * NE]) {
#pragma unroll
    for (int cix = 0; cix < 3 * NE; cix++) {
        const double* t = P->sd[g][cix];
        const double zr = fma(fr.c, t[0], -fma(fr.al, t[3], fr.be * t[5]));
        const double zi = fma(fr.c, t[1], fma(fr.al, t[2], fr.be * t[4]));
        const double a0 = P->dabs[g][cix];
        dsq[cix] = fr.ident ? a0 * a0 : fma(zr, zr, zi * zi);
    }
}

// Doppler shift and Gaussian factor of this lane's point for each distinct
// beam (direction, radius), into its LDS column: one 16-byte pair per
// direction at col2[d * 64] (col2 = the lane's base, stride 64 pairs).  The
// six-beam configuration has 6 directions for its 24 beams, so the column
// stays small enough for two waves per SIMD.
__device__ __forceinline__ void s3_spin_beams(const S3Plan* __restrict__ P, const double (&r)[3],
                                              const double (&v)[3], double2* __restrict__ col2) {
    const double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    for (int d = 0; d < P->nd; d++) {
        const double kx = P->dk[d][0], ky = P->dk[d][1], kz = P->dk[d][2];
        const double dop = -P->k * (kx * v[0] + ky * v[1] + kz * v[2]) / (2 * M_PI);
        double gauss = 1.0;
        if (P->dw[d] > 0.0) {
            const double rpar = r[0] * kx + r[1] * ky + r[2] * kz;
            const double rp2 = fmax(rr - rpar * rpar, 0.0);
            gauss = exp(-2.0 * rp2 / (P->dw[d] * P->dw[d]));
        }
        col2[d * 64] = make_double2(dop, gauss);
    }
}

// Rates of ground state g: unscaled R~_ge (Rt), k-weighted sum (fk) and
// saturation total (st); the beam sums in class order.
template <int NE>
__device__ __forceinline__ void s3_spin_rates(const S3Plan* __restrict__ P, const S3SpinFrame& fr, int g,
                                              double Esg, const double2* __restrict__ col2, double (&BR)[NE],
                                              double (&Rt)[NE], double (&fk)[3], double& st) {
    // L = (Gamma/2)^2 / (de^2 + (Gamma/2)^2) = 1 / (1 + x^2), x = (f - om) 4 pi / Gamma;
    // the difference f - om is taken first, as the reference does
    const double xs = 4.0 * M_PI / P->Gamma;
    double om[NE], A[3][NE];
#pragma unroll
    for (int e = 0; e < NE; e++) {
        om[e] = fr.Ese[e] - Esg;
        double d2[3], dq[3];
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const double* t = P->sd[g][e * 3 + q];
            const double zr = fma(fr.c, t[0], -fma(fr.al, t[3], fr.be * t[5]));
            const double zi = fma(fr.c, t[1], fma(fr.al, t[2], fr.be * t[4]));
            const double a0 = P->dabs[g][e * 3 + q];
            const double d = fr.ident ? a0 * a0 : fma(zr, zr, zi * zi);
            dq[q] = d;
            d2[q] = d < 1e-15 ? 0.0 : d;
        }
        BR[e] = ((dq[0] + dq[1]) + dq[2]) * P->inv_tot[e];
#pragma unroll
        for (int ql = 0; ql < 3; ql++) A[ql][e] = (fr.pw[ql][0] * d2[0] + fr.pw[ql][1] * d2[1]) + fr.pw[ql][2] * d2[2];
        Rt[e] = 0.0;
    }
    fk[0] = fk[1] = fk[2] = 0.0;
    st = 0.0;
#pragma unroll
    for (int ql = 0; ql < 3; ql++) {
#pragma unroll 2
        for (int t = P->cls[ql]; t < P->cls[ql + 1]; t++) {
            const double2 dg = col2[P->pdir[t] * 64];
            const double2 fs = make_double2(P->pf[t] + dg.x, P->ps0[t] * dg.y);
            double sb = 0.0;
#pragma unroll
            for (int e = 0; e < NE; e++) {
                const double x = (fs.x - om[e]) * xs;
                const double y = fma(x, x, 1.0);
                double L = __builtin_amdgcn_rcp(y);
                L = fma(L, fma(-y, L, 1.0), L);
                const double rs = L * A[ql][e];
                sb = fma(fs.y, rs, sb);
                Rt[e] = fma(rs, fs.y, Rt[e]);
            }
            st += sb;
            fk[0] = fma(P->pk[t][0], sb, fk[0]);
            fk[1] = fma(P->pk[t][1], sb, fk[1]);
            fk[2] = fma(P->pk[t][2], sb, fk[2]);
        }
    }
}

// The reference's generator solve at one point (simulator_3d.py:318-365):
// 12 x 12 generator with its last row replaced by ones, elimination with
// partial pivoting (first maximum), clamp, normalise.  Only for points the
// reduction does not cover; private arrays, one lane.
template <int NG, int NE>
__device__ __attribute__((noinline)) void s3_spin_dense(const S3Plan* __restrict__ P, double r0, double r1,
                                                        double r2, double Bg, const double2* __restrict__ col2,
                                                        double* F, double* Rsc, double* __restrict__ pop) {
    // the frame is rebuilt here from (r, Bg): passing the caller's by
    // address would keep it in scratch memory on the fast path too.  Rolled
    // loops throughout: this path is rare, and its registers count towards
    // the kernel's.
    const double r[3] = {r0, r1, r2};
    S3SpinFrame fr;
    s3_spin_frame<NE>(P, r, Bg, fr);
    const double G = P->Gamma, gh = G / 2.0, gh2 = gh * gh;
    const double G2 = G / (2 * M_PI);
    double Rt[NG][NE], FK[NG][3], BR[NE][NG], M[NG][NG + 1], x[NG];
#pragma unroll 1
    for (int g = 0; g < NG; g++) {
        const double Esg = P->E[g] + fr.Bm * P->lam0[g] * G2;
        double d2[NE][3], A[3][NE], om[NE];
#pragma unroll 1
        for (int e = 0; e < NE; e++) {
            om[e] = fr.Ese[e] - Esg;
            double bn = 0.0;
#pragma unroll 1
            for (int q = 0; q < 3; q++) {
                const double* t = P->sd[g][e * 3 + q];
                const double zr = fma(fr.c, t[0], -fma(fr.al, t[3], fr.be * t[5]));
                const double zi = fma(fr.c, t[1], fma(fr.al, t[2], fr.be * t[4]));
                const double a0 = P->dabs[g][e * 3 + q];
                const double d = fr.ident ? a0 * a0 : fma(zr, zr, zi * zi);
                bn = q ? bn + d : d;
                d2[e][q] = d < 1e-15 ? 0.0 : d;
            }
            BR[e][g] = bn * P->inv_tot[e];
#pragma unroll 1
            for (int ql = 0; ql < 3; ql++)
                A[ql][e] = (fr.pw[ql][0] * d2[e][0] + fr.pw[ql][1] * d2[e][1]) + fr.pw[ql][2] * d2[e][2];
            Rt[g][e] = 0.0;
        }
        double st = 0.0, fk[3] = {0.0, 0.0, 0.0};
#pragma unroll 1
        for (int ql = 0; ql < 3; ql++)
#pragma unroll 1
            for (int t = P->cls[ql]; t < P->cls[ql + 1]; t++) {
                const double2 dg = col2[P->pdir[t] * 64];
                const double2 fs = make_double2(P->pf[t] + dg.x, P->ps0[t] * dg.y);
                double sb = 0.0;
#pragma unroll 1
                for (int e = 0; e < NE; e++) {
                    const double de = (fs.x - om[e]) * (2 * M_PI);
                    const double L = gh2 / (de * de + gh2);
                    const double rs = L * A[ql][e];
                    sb = fma(fs.y, rs, sb);
                    Rt[g][e] = fma(rs, fs.y, Rt[g][e]);
                }
                st += sb;
                fk[0] = fma(P->pk[t][0], sb, fk[0]);
                fk[1] = fma(P->pk[t][1], sb, fk[1]);
                fk[2] = fma(P->pk[t][2], sb, fk[2]);
            }
        const double sc = gh * P->Geff / (1.0 + st);
#pragma unroll 1
        for (int e = 0; e < NE; e++) Rt[g][e] *= sc;
        FK[g][0] = sc * fk[0];
        FK[g][1] = sc * fk[1];
        FK[g][2] = sc * fk[2];
    }
#pragma unroll 1
    for (int g = 0; g < NG; g++) {
        double sg = Rt[g][0];
        for (int e = 1; e < NE; e++) sg += Rt[g][e];
#pragma unroll 1
        for (int k2 = 0; k2 < NG; k2++) {
            double m = (k2 == g) ? -sg : 0.0;
            for (int e = 0; e < NE; e++) m += Rt[k2][e] * BR[e][g];
            M[g][k2] = (g == NG - 1) ? 1.0 : m;
        }
        M[g][NG] = (g == NG - 1) ? 1.0 : 0.0;
    }
    bool sing = false;
#pragma unroll 1
    for (int cc = 0; cc < NG; cc++) {
        double best = -1.0;
        int piv = cc;
        for (int i = cc; i < NG; i++) {
            const double a = fabs(M[i][cc]);
            if (a > best) {
                best = a;
                piv = i;
            }
        }
        if (best == 0.0) sing = true;
        if (piv != cc)
            for (int j = 0; j <= NG; j++) {
                const double t = M[cc][j];
                M[cc][j] = M[piv][j];
                M[piv][j] = t;
            }
        if (M[cc][cc] != 0.0)
            for (int i = cc + 1; i < NG; i++) {
                const double l = M[i][cc] / M[cc][cc];
                for (int j = cc + 1; j <= NG; j++) M[i][j] -= l * M[cc][j];
                M[i][cc] = 0.0;
            }
    }
    if (sing) {
        for (int i = 0; i < NG; i++) x[i] = 1.0 / NG;
    } else {
#pragma unroll 1
        for (int i = NG - 1; i >= 0; i--) {
            double acc = M[i][NG];
            for (int j = i + 1; j < NG; j++) acc -= M[i][j] * x[j];
            x[i] = acc / M[i][i];
        }
        for (int i = 0; i < NG; i++) x[i] = fmax(x[i], 0.0);
        const double ps = np_sum<NG>(x);
        for (int i = 0; i < NG; i++) x[i] = ps > 0.0 ? x[i] / ps : 1.0 / NG;
    }
    double fx = 0.0, fy = 0.0, fz = 0.0, rs = 0.0;
#pragma unroll 1
    for (int g = 0; g < NG; g++) {
        fx += FK[g][0] * x[g];
        fy += FK[g][1] * x[g];
        fz += FK[g][2] * x[g];
        double sg = Rt[g][0];
        for (int e = 1; e < NE; e++) sg += Rt[g][e];
        rs += sg * x[g];
        if (pop) pop[g] = x[g];
    }
    const double hk = P->hbar * P->k;
    F[0] = hk * fx;
    F[1] = hk * fy;
    F[2] = hk * fz;
    *Rsc = rs;
}

struct S3FR {
    double F0, F1, F2, R;
};

// sum over the L lanes of a group (L a power of two <= 16): butterfly, so
// every lane of the group ends with the same bits
template <int L>
__device__ __forceinline__ double s3_lsum(double x) {
#pragma unroll
    for (int m = L / 2; m >= 1; m >>= 1) x += __shfl_xor(x, m, L);
    return x;
}

// Force and scattering rate at one point; `col` is the lane's LDS column
// (stride 64, one pair per beam direction); `pop` (global, NG; L = 1 only)
// receives the populations when not null.  With L > 1 the point belongs to
// a group of L lanes, lane `sub` taking ground states sub, sub + L, ...; the
// per-ground sums are added across the group and every lane returns the
// same result.  Not inlined: the trajectory kernel would otherwise keep
// step-invariant plan values live across its step loop and run out of
// registers.
template <int NG, int NE, int L = 1>
__device__ __attribute__((noinline)) S3FR s3_spin_point(const S3Plan* __restrict__ P, double r0, double r1,
                                                        double r2, double v0, double v1, double v2, double Bg,
                                                        double2* __restrict__ col, double* __restrict__ pop,
                                                        int sub = 0) {
    const double r[3] = {r0, r1, r2}, v[3] = {v0, v1, v2};
    double F[3], Rsc;
    S3SpinFrame fr;
    s3_spin_frame<NE>(P, r, Bg, fr);
    s3_spin_beams(P, r, v, col);
    const double G = P->Gamma, gh = G / 2.0;
    const double G2 = G / (2 * M_PI);
    double K[NE][NE], PF[NE][3], ZF[NE], UF[NE];
#pragma unroll
    for (int e = 0; e < NE; e++) {
#pragma unroll
        for (int f = 0; f < NE; f++) K[e][f] = 0.0;
        PF[e][0] = PF[e][1] = PF[e][2] = 0.0;
        ZF[e] = UF[e] = 0.0;
    }
    bool ok = true;
#pragma unroll 1
    for (int g = sub; g < NG; g += L) {
        double BR[NE], Rt[NE], fk[3], st;
        s3_spin_rates<NE>(P, fr, g, P->E[g] + fr.Bm * P->lam0[g] * G2, col, BR, Rt, fk, st);
        double Rg = Rt[0];
#pragma unroll
        for (int e = 1; e < NE; e++) Rg += Rt[e];
        ok = ok && (Rg > 0.0);
        const double rR = s3_rcp(Rg);
        const double sc = gh * P->Geff / (1.0 + st);
        const double qg = rR / sc;  // 1 / D_g
#pragma unroll
        for (int e = 0; e < NE; e++) {
            const double a = Rt[e] * rR;
#pragma unroll
            for (int f = 0; f < NE; f++)
                if (f != e) K[e][f] = fma(a, BR[f], K[e][f]);
        }
#pragma unroll
        for (int f = 0; f < NE; f++) {
            const double bq = BR[f] * rR;
            PF[f][0] = fma(bq, fk[0], PF[f][0]);
            PF[f][1] = fma(bq, fk[1], PF[f][1]);
            PF[f][2] = fma(bq, fk[2], PF[f][2]);
            ZF[f] = fma(BR[f], qg, ZF[f]);
            UF[f] += BR[f];
        }
        if (pop) pop[g] = qg;
    }
    if constexpr (L > 1) {
#pragma unroll
        for (int e = 0; e < NE; e++) {
#pragma unroll
            for (int f = 0; f < NE; f++)
                if (f != e) K[e][f] = s3_lsum<L>(K[e][f]);
            PF[e][0] = s3_lsum<L>(PF[e][0]);
            PF[e][1] = s3_lsum<L>(PF[e][1]);
            PF[e][2] = s3_lsum<L>(PF[e][2]);
            ZF[e] = s3_lsum<L>(ZF[e]);
            UF[e] = s3_lsum<L>(UF[e]);
        }
        int okl = ok ? 1 : 0;
#pragma unroll
        for (int m = L / 2; m >= 1; m >>= 1) okl &= __shfl_xor(okl, m, L);
        ok = okl != 0;
    }
    // GTH state reduction of the excited chain (off-diagonal K only)
#pragma unroll
    for (int m = NE - 1; m >= 1; m--) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < m; j++) s += K[j][m];
        ok = ok && (s > 0.0);
        const double rs = s3_rcp(s);
#pragma unroll
        for (int i = 0; i < m; i++) K[m][i] *= rs;
#pragma unroll
        for (int i = 0; i < m; i++)
#pragma unroll
            for (int j = 0; j < m; j++)
                if (i != j) K[j][i] = fma(K[m][i], K[j][m], K[j][i]);
    }
    if (!ok) {
        double Fd[3], Rd;  // temporaries: &F, &Rsc would keep them in scratch memory
        s3_spin_dense<NG, NE>(P, r[0], r[1], r[2], Bg, col, Fd, &Rd, pop);
        return S3FR{Fd[0], Fd[1], Fd[2], Rd};
    }
    double y[NE];
    y[0] = 1.0;
#pragma unroll
    for (int j = 1; j < NE; j++) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < j; i++) acc = fma(y[i], K[j][i], acc);
        y[j] = acc;
    }
    double Z = 0.0, fx = 0.0, fy = 0.0, fz = 0.0, rs = 0.0;
#pragma unroll
    for (int f = 0; f < NE; f++) {
        Z = fma(y[f], ZF[f], Z);
        fx = fma(y[f], PF[f][0], fx);
        fy = fma(y[f], PF[f][1], fy);
        fz = fma(y[f], PF[f][2], fz);
        rs = fma(y[f], UF[f], rs);
    }
    const double rZ = 1.0 / Z;
    const double hk = P->hbar * P->k;
    F[0] = hk * fx * rZ;
    F[1] = hk * fy * rZ;
    F[2] = hk * fz * rZ;
    Rsc = rs * rZ;
    if (pop) {
#pragma unroll 1
        for (int g = 0; g < NG; g++) {
            double dsq[3 * NE];
            s3_spin_dsq<NE>(P, fr, g, dsq);
            double u = 0.0;
#pragma unroll
            for (int f = 0; f < NE; f++)
                u = fma(y[f], ((dsq[f * 3] + dsq[f * 3 + 1]) + dsq[f * 3 + 2]) * P->inv_tot[f], u);
            pop[g] = u * pop[g] * rZ;
        }
    }
    return S3FR{F[0], F[1], F[2], Rsc};
}

template <int NG, int NE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_force3d_spin(const S3Plan* __restrict__ P, long n,
                                                     const double* __restrict__ r_in, const double* __restrict__ v_in,
                                                     double Bg, double* __restrict__ F_out, double* __restrict__ R_out,
                                                     double* __restrict__ pop_out) {
    extern __shared__ double2 s3dyn[];
    const long i = (long)blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    double r[3], v[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r[a] = r_in[i * 3 + a];
        v[a] = v_in[i * 3 + a];
    }
    const S3FR o = s3_spin_point<NG, NE>(P, r[0], r[1], r[2], v[0], v[1], v[2], Bg, s3dyn + threadIdx.x,
                                         pop_out ? pop_out + i * NG : nullptr);
    F_out[i * 3] = o.F0;
    F_out[i * 3 + 1] = o.F1;
    F_out[i * 3 + 2] = o.F2;
    if (R_out) R_out[i] = o.R;
}

// Leapfrog of k_leapfrog3d, one particle per lane (same step arithmetic).
template <int NG, int NE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_leapfrog3d_spin(const S3Plan* __restrict__ P, long n,
                                                        const double* __restrict__ r0, const double* __restrict__ v0,
                                                        double Bg, double dt, int n_steps, double r_escape,
                                                        double* __restrict__ r_traj, double* __restrict__ v_traj,
                                                        double* __restrict__ r_fin, double* __restrict__ v_fin,
                                                        int* __restrict__ n_act) {
    extern __shared__ double2 s3dyn[];
    const long i = (long)blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    double2* col = s3dyn + threadIdx.x;
    const double m = P->mass;
    double r[3], vh[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        r[a] = r0[i * 3 + a];
        vh[a] = v0[i * 3 + a];
    }
    if (r_traj)
#pragma unroll
        for (int a = 0; a < 3; a++) {
            r_traj[(i * n_steps) * 3 + a] = r[a];
            v_traj[(i * n_steps) * 3 + a] = vh[a];
        }
    // step 0 is the initial half kick; one force call site for all steps
    int nact = n_steps;
    for (int j = 0; j < n_steps; j++) {
        if (j > 0) {
            double rn[3];
#pragma unroll
            for (int a = 0; a < 3; a++) rn[a] = r[a] + dt * vh[a];
            const double nr = sqrt(rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2]);
#pragma unroll
            for (int a = 0; a < 3; a++) r[a] = rn[a];
            if (nr > r_escape) {
                nact = j;
                if (r_traj)
                    for (int jj = j; jj < n_steps; jj++)
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            r_traj[(i * n_steps + jj) * 3 + a] = r[a];
                            v_traj[(i * n_steps + jj) * 3 + a] = vh[a];
                        }
                break;
            }
        }
        const S3FR o = s3_spin_point<NG, NE>(P, r[0], r[1], r[2], vh[0], vh[1], vh[2], Bg, col, nullptr);
        const double F[3] = {o.F0, o.F1, o.F2};
        if (j == 0) {
#pragma unroll
            for (int a = 0; a < 3; a++) vh[a] = vh[a] + 0.5 * dt * F[a] / m;
            continue;
        }
        double vs[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double vn = vh[a] + dt * F[a] / m;
            vs[a] = 0.5 * (vh[a] + vn);
            vh[a] = vn;
        }
        if (r_traj)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                r_traj[(i * n_steps + j) * 3 + a] = r[a];
                v_traj[(i * n_steps + j) * 3 + a] = vs[a];
            }
    }
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (r_fin) r_fin[i * 3 + a] = r[a];
        if (v_fin) v_fin[i * 3 + a] = vh[a];
    }
    if (n_act) n_act[i] = nact;
}

// One segment [j0, j1) of the leapfrog for the particles listed in idx[0,
// A) (idx = NULL: particles 0..A-1), L lanes per particle.  A particle's
// state (r, v_half, n_actual) lives in r_fin, v_fin, n_act between
// segments; segment j0 = 0 starts from r0, v0 with the initial half kick.
// The steps are those of k_leapfrog3d_spin; with L = 1 the results are
// bit-identical to it, with L > 1 the per-ground sums are added in a
// different order.
template <int NG, int NE, int L>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_leapfrog3d_spin_seg(
    const S3Plan* __restrict__ P, const int* __restrict__ idx, long A, const double* __restrict__ r0,
    const double* __restrict__ v0, double Bg, double dt, int n_steps, int j0, int j1, double r_escape,
    double* __restrict__ r_traj, double* __restrict__ v_traj, double* __restrict__ r_fin,
    double* __restrict__ v_fin, int* __restrict__ n_act) {
    extern __shared__ double2 s3dyn[];
    const long slot = (long)blockIdx.x * (64 / L) + threadIdx.x / L;
    const int sub = (int)(threadIdx.x % L);
    if (slot >= A) return;  // a whole group leaves together
    const long i = idx ? (long)idx[slot] : slot;
    double2* col = s3dyn + threadIdx.x;
    const double m = P->mass;
    double r[3], vh[3];
    if (j0 == 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            r[a] = r0[i * 3 + a];
            vh[a] = v0[i * 3 + a];
        }
        if (r_traj && sub == 0)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                r_traj[(i * n_steps) * 3 + a] = r[a];
                v_traj[(i * n_steps) * 3 + a] = vh[a];
            }
    } else {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            r[a] = r_fin[i * 3 + a];
            vh[a] = v_fin[i * 3 + a];
        }
    }
    int nact = n_steps;
    for (int j = j0; j < j1; j++) {
        if (j > 0) {
            double rn[3];
#pragma unroll
            for (int a = 0; a < 3; a++) rn[a] = r[a] + dt * vh[a];
            const double nr = sqrt(rn[0] * rn[0] + rn[1] * rn[1] + rn[2] * rn[2]);
#pragma unroll
            for (int a = 0; a < 3; a++) r[a] = rn[a];
            if (nr > r_escape) {
                nact = j;
                if (r_traj && sub == 0)
                    for (int jj = j; jj < n_steps; jj++)
#pragma unroll
                        for (int a = 0; a < 3; a++) {
                            r_traj[(i * n_steps + jj) * 3 + a] = r[a];
                            v_traj[(i * n_steps + jj) * 3 + a] = vh[a];
                        }
                break;
            }
        }
        const S3FR o = s3_spin_point<NG, NE, L>(P, r[0], r[1], r[2], vh[0], vh[1], vh[2], Bg, col, nullptr, sub);
        const double F[3] = {o.F0, o.F1, o.F2};
        if (j == 0) {
#pragma unroll
            for (int a = 0; a < 3; a++) vh[a] = vh[a] + 0.5 * dt * F[a] / m;
            continue;
        }
        double vs[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double vn = vh[a] + dt * F[a] / m;
            vs[a] = 0.5 * (vh[a] + vn);
            vh[a] = vn;
        }
        if (r_traj && sub == 0)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                r_traj[(i * n_steps + j) * 3 + a] = r[a];
                v_traj[(i * n_steps + j) * 3 + a] = vs[a];
            }
    }
    if (sub == 0) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            r_fin[i * 3 + a] = r[a];
            v_fin[i * 3 + a] = vh[a];
        }
        n_act[i] = nact;
    }
}

// The particles of idx_in[0, A) (idx_in = NULL: 0..A-1) with n_act =
// n_steps, in order, into idx_out; their number into *cnt.  One block of
// 1024 threads: ballot prefix per wavefront, wavefront offsets in LDS.
__global__ __launch_bounds__(1024) void k_s3_compact(const int* __restrict__ idx_in, long A,
                                                     const int* __restrict__ n_act, int n_steps,
                                                     int* __restrict__ idx_out, int* __restrict__ cnt) {
    __shared__ int wsum[16];
    __shared__ int base;
    const int t = threadIdx.x, w = t >> 6, l = t & 63;
    if (t == 0) base = 0;
    __syncthreads();
    for (long c = 0; c < A; c += 1024) {
        const long k = c + t;
        int id = 0, keep = 0;
        if (k < A) {
            id = idx_in ? idx_in[k] : (int)k;
            keep = n_act[id] == n_steps;
        }
        const unsigned long long bal = __ballot(keep);
        const int pre = __popcll(bal & ((1ull << l) - 1ull));
        if (l == 0) wsum[w] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int q = 0; q < w; q++) off += wsum[q];
        if (keep) idx_out[off + pre] = id;
        __syncthreads();
        if (t == 0) {
            int tot = 0;
            for (int q = 0; q < 16; q++) tot += wsum[q];
            base += tot;
        }
        __syncthreads();
    }
    if (t == 0) *cnt = base;
}

#define S3_SIZE_LIST(X) X(12, 4) X(1, 1) X(2, 2)

extern "C" {

int mm_force3d_plan_create(mm_force3d_plan** out, int ng, int ne, int nb, const double* energies,
                           const double* zeeman_re, const double* zeeman_im, const double* zdiag_e,
                           const double* dabs, const double* beam_k, const double* beam_freq, const int* beam_q,
                           const double* beam_s0, const double* beam_w, double Gamma, double k, double Geff,
                           double hbar, double mass) {
    mm_clear_error();
    if (!out) return mm_fail(MM_EINVAL, "force3d_plan_create: null output");
    *out = nullptr;
    bool ok = false;
#define X(G, E) ok = ok || (ng == G && ne == E);
    S3_SIZE_LIST(X)
#undef X
    if (!ok) return mm_fail(MM_EUNSUPPORTED, "force3d: no kernel for %d ground + %d excited states", ng, ne);
    if (nb < 0 || nb > S3_MAXB) return mm_fail(MM_EINVAL, "force3d: %d beams (0..%d)", nb, S3_MAXB);
    if (!energies || !zeeman_re || !zeeman_im || !zdiag_e || !dabs || (nb > 0 && (!beam_k || !beam_freq ||
        !beam_q || !beam_s0 || !beam_w)))
        return mm_fail(MM_EINVAL, "force3d_plan_create: null table");
    if (!(Gamma > 0.0) || !(mass > 0.0)) return mm_fail(MM_EINVAL, "force3d: Gamma and mass must be positive");
    for (int b = 0; b < nb; b++)
        if (beam_q[b] < 0 || beam_q[b] > 2) return mm_fail(MM_EINVAL, "force3d: beam %d polarisation %d", b, beam_q[b]);
    mm_force3d_plan* p = new mm_force3d_plan();
    S3Plan& h = p->host;
    h.ng = ng;
    h.ne = ne;
    h.nb = nb;
    h.Gamma = Gamma;
    h.k = k;
    h.Geff = Geff;
    h.hbar = hbar;
    h.mass = mass;
    for (int i = 0; i < ng + ne; i++) h.E[i] = energies[i];
    for (int a = 0; a < 3; a++)
        for (int i = 0; i < ng; i++)
            for (int j = 0; j < ng; j++) {
                h.zr[a][i][j] = zeeman_re[(a * ng + i) * ng + j];
                h.zi[a][i][j] = zeeman_im[(a * ng + i) * ng + j];
            }
    for (int e = 0; e < ne; e++) h.zde[e] = zdiag_e[e];
    for (int g = 0; g < ng; g++)
        for (int c = 0; c < 3 * ne; c++) h.dabs[g][c] = dabs[g * 3 * ne + c];
    for (int b = 0; b < nb; b++) {
        for (int a = 0; a < 3; a++) h.bk[b][a] = beam_k[b * 3 + a];
        h.bfreq[b] = beam_freq[b];
        h.bq[b] = beam_q[b];
        h.bs0[b] = beam_s0[b];
        h.bw[b] = beam_w[b];
    }
    // round-robin schedule over m = ng rounded up to even players (player m-1 fixed)
    const int m = (ng + 1) / 2 * 2;
    h.n_rounds = ng > 1 ? m - 1 : 0;
    h.n_slots = m / 2;
    for (int rd = 0; rd < h.n_rounds; rd++)
        for (int sl = 0; sl < h.n_slots; sl++) {
            int a, b;
            if (sl == 0) {
                a = m - 1;
                b = rd;
            } else {
                a = (rd + sl) % (m - 1);
                b = (rd - sl + (m - 1)) % (m - 1);
            }
            if (a >= ng || b >= ng) a = b = -1;
            h.pairs[rd][sl][0] = a < b ? a : b;
            h.pairs[rd][sl][1] = a < b ? b : a;
        }
    // beams grouped by lab polarisation for the spin-basis kernels (stable)
    h.spin = 0;
    {
        int t = 0;
        for (int q = 0; q < 3; q++) {
            h.cls[q] = t;
            for (int b = 0; b < nb; b++)
                if (h.bq[b] == q) h.perm[t++] = b;
        }
        h.cls[3] = t;
        h.nd = 0;
        for (int u = 0; u < nb; u++) {
            const int b = h.perm[u];
            for (int a = 0; a < 3; a++) h.pk[u][a] = h.bk[b][a];
            h.pf[u] = h.bfreq[b];
            h.ps0[u] = h.bs0[b];
            h.pwr[u] = h.bw[b];
            int d = 0;
            while (d < h.nd && !(h.dk[d][0] == h.pk[u][0] && h.dk[d][1] == h.pk[u][1] && h.dk[d][2] == h.pk[u][2] &&
                                 h.dw[d] == h.pwr[u]))
                d++;
            if (d == h.nd) {
                for (int a = 0; a < 3; a++) h.dk[d][a] = h.pk[u][a];
                h.dw[d] = h.pwr[u];
                h.nd++;
            }
            h.pdir[u] = d;
        }
    }
    if (hipMalloc(&p->dev, sizeof(S3Plan)) != hipSuccess) {
        delete p;
        return mm_hip_fail("force3d_plan_create: hipMalloc");
    }
    if (hipMemcpy(p->dev, &h, sizeof(S3Plan), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p->dev);
        delete p;
        return mm_hip_fail("force3d_plan_create: hipMemcpy");
    }
    *out = p;
    return MM_OK;
}

int mm_force3d_plan_set_spin_basis(mm_force3d_plan* p, const double* lam0, const double* sd, const double* inv_tot) {
    mm_clear_error();
    if (!p) return mm_fail(MM_EINVAL, "force3d_plan_set_spin_basis: null plan");
    S3Plan& h = p->host;
    if (!lam0) {
        h.spin = 0;
    } else {
        if (!sd || !inv_tot) return mm_fail(MM_EINVAL, "force3d_plan_set_spin_basis: null table");
        for (int g = 0; g < h.ng; g++) {
            h.lam0[g] = lam0[g];
            if (g > 0 && !(lam0[g] >= lam0[g - 1]))
                return mm_fail(MM_EINVAL, "force3d_plan_set_spin_basis: eigenvalues must ascend");
            for (int c = 0; c < 3 * h.ne; c++)
                for (int k = 0; k < 6; k++) h.sd[g][c][k] = sd[(g * 3 * h.ne + c) * 6 + k];
        }
        for (int e = 0; e < h.ne; e++) {
            if (!(inv_tot[e] >= 0.0)) return mm_fail(MM_EINVAL, "force3d_plan_set_spin_basis: inv_tot < 0");
            h.inv_tot[e] = inv_tot[e];
        }
        h.spin = 1;
    }
    if (hipMemcpy(p->dev, &h, sizeof(S3Plan), hipMemcpyHostToDevice) != hipSuccess)
        return mm_hip_fail("force3d_plan_set_spin_basis: hipMemcpy");
    return MM_OK;
}

int mm_force3d_plan_destroy(mm_force3d_plan* p) {
    mm_clear_error();
    if (!p) return MM_OK;
    (void)hipFree(p->dev);
    delete p;
    return MM_OK;
}

int mm_force3d_points(const mm_force3d_plan* p, long n, const double* r, const double* v, double B_gradient,
                      double* F, double* R, double* pop, void* stream) {
    mm_clear_error();
    if (!p) return mm_fail(MM_EINVAL, "force3d_points: null plan");
    if (n < 0) return mm_fail(MM_EINVAL, "force3d_points: negative size");
    if (n == 0) return MM_OK;
    if (!r || !v || !F) return mm_fail(MM_EINVAL, "force3d_points: null pointer");
    const unsigned grid = (unsigned)((n + S3_PTS - 1) / S3_PTS);
    const unsigned grid_s = (unsigned)((n + 63) / 64);
    const size_t lds_s = (size_t)2 * p->host.nd * 64 * sizeof(double);
    const int ng = p->host.ng, ne = p->host.ne;
#define X(G, E)                                                                                              \
    if (ng == G && ne == E) {                                                                                \
        if (p->host.spin) {                                                                                  \
            hipLaunchKernelGGL((k_force3d_spin<G, E>), dim3(grid_s), dim3(64), lds_s, (hipStream_t)stream,     \
                               p->dev, n, r, v, B_gradient, F, R, pop);                                      \
            return mm_check_launch("k_force3d_spin");                                                        \
        }                                                                                                    \
        hipLaunchKernelGGL((k_force3d<G, E>), dim3(grid), dim3(64), 0, (hipStream_t)stream, p->dev, n, r, v, \
                           B_gradient, F, R, pop);                                                           \
        return mm_check_launch("k_force3d");                                                                 \
    }
    S3_SIZE_LIST(X)
#undef X
    return mm_fail(MM_EUNSUPPORTED, "force3d_points: no kernel");
}

int mm_force3d_leapfrog(const mm_force3d_plan* p, long n, const double* r0, const double* v0, double B_gradient,
                        double dt, int n_steps, double r_escape, double* r_traj, double* v_traj, double* r_final,
                        double* v_final, int* n_actual, void* stream) {
    mm_clear_error();
    if (!p) return mm_fail(MM_EINVAL, "force3d_leapfrog: null plan");
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "force3d_leapfrog: n >= 0 and n_steps >= 1");
    if (!(dt > 0.0)) return mm_fail(MM_EINVAL, "force3d_leapfrog: dt must be positive");
    if (n == 0) return MM_OK;
    if (!r0 || !v0) return mm_fail(MM_EINVAL, "force3d_leapfrog: null pointer");
    if ((r_traj == nullptr) != (v_traj == nullptr))
        return mm_fail(MM_EINVAL, "force3d_leapfrog: r_traj and v_traj go together");
    const unsigned grid = (unsigned)((n + S3_PTS - 1) / S3_PTS);
    const unsigned grid_s = (unsigned)((n + 63) / 64);
    const size_t lds_s = (size_t)2 * p->host.nd * 64 * sizeof(double);
    const int ng = p->host.ng, ne = p->host.ne;
#define X(G, E)                                                                                                 \
    if (ng == G && ne == E) {                                                                                   \
        if (p->host.spin) {                                                                                     \
            hipLaunchKernelGGL((k_leapfrog3d_spin<G, E>), dim3(grid_s), dim3(64), lds_s, (hipStream_t)stream,     \
                               p->dev, n, r0, v0, B_gradient, dt, n_steps, r_escape, r_traj, v_traj, r_final,   \
                               v_final, n_actual);                                                              \
            return mm_check_launch("k_leapfrog3d_spin");                                                        \
        }                                                                                                       \
        hipLaunchKernelGGL((k_leapfrog3d<G, E>), dim3(grid), dim3(64), 0, (hipStream_t)stream, p->dev, n, r0, v0, \
                           B_gradient, dt, n_steps, r_escape, r_traj, v_traj, r_final, v_final, n_actual);       \
        return mm_check_launch("k_leapfrog3d");                                                                 \
    }
    S3_SIZE_LIST(X)
#undef X
    return mm_fail(MM_EUNSUPPORTED, "force3d_leapfrog: no kernel");
}

long mm_force3d_leapfrog_workspace_bytes(long n) {
    if (n < 0) return 0;
    const long ib = (n * (long)sizeof(int) + 255) / 256 * 256;
    return 256 + 2 * ib;
}

int mm_force3d_leapfrog_ws(const mm_force3d_plan* p, long n, const double* r0, const double* v0, double B_gradient,
                           double dt, int n_steps, double r_escape, double* r_traj, double* v_traj, double* r_final,
                           double* v_final, int* n_actual, int seg_steps, void* ws, long ws_bytes, void* stream) {
    mm_clear_error();
    if (!p) return mm_fail(MM_EINVAL, "force3d_leapfrog_ws: null plan");
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "force3d_leapfrog_ws: n >= 0 and n_steps >= 1");
    if (n > 0x7fffffffL) return mm_fail(MM_EINVAL, "force3d_leapfrog_ws: at most 2^31 - 1 particles");
    if (!(dt > 0.0)) return mm_fail(MM_EINVAL, "force3d_leapfrog_ws: dt must be positive");
    if (seg_steps < 0) return mm_fail(MM_EINVAL, "force3d_leapfrog_ws: seg_steps >= 0");
    if (n == 0) return MM_OK;
    if (!r0 || !v0 || !r_final || !v_final || !n_actual)
        return mm_fail(MM_EINVAL, "force3d_leapfrog_ws: r0, v0, r_final, v_final and n_actual are required");
    if ((r_traj == nullptr) != (v_traj == nullptr))
        return mm_fail(MM_EINVAL, "force3d_leapfrog_ws: r_traj and v_traj go together");
    if (!p->host.spin)
        return mm_force3d_leapfrog(p, n, r0, v0, B_gradient, dt, n_steps, r_escape, r_traj, v_traj, r_final,
                                   v_final, n_actual, stream);
    if (!ws || ws_bytes < mm_force3d_leapfrog_workspace_bytes(n) || ((uintptr_t)ws & 255))
        return mm_fail(MM_EINVAL, "force3d_leapfrog_ws: workspace of mm_force3d_leapfrog_workspace_bytes(n) "
                                  "bytes, 256-byte aligned, required");
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
        return mm_hip_fail("force3d_leapfrog_ws: device query");
    const long nsimd = 4l * ncu;
    const int seg = seg_steps > 0 ? seg_steps : 256;
    hipStream_t s = (hipStream_t)stream;
    int* cnt = (int*)ws;
    int* ia = (int*)((char*)ws + 256);
    int* ib = (int*)((char*)ws + 256 + (mm_force3d_leapfrog_workspace_bytes(n) - 256) / 2);
    const size_t lds_s = (size_t)2 * p->host.nd * 64 * sizeof(double);
    const int ng = p->host.ng, ne = p->host.ne;
    long A = n;
    const int* idx = nullptr;
    int* out = ia;
    for (int j0 = 0; j0 < n_steps && A > 0; j0 += seg) {
        const int j1 = (n_steps - j0 > seg) ? j0 + seg : n_steps;
        // lanes per particle: the most that keep the alive particles at
        // <= 2 wavefronts per SIMD
        const int L = (A * 16 <= 128 * nsimd) ? 16 : (A * 4 <= 128 * nsimd) ? 4 : 1;
        const unsigned grid = (unsigned)((A * L + 63) / 64);
        bool done = false;
#define X(G, E)                                                                                                  \
        if (!done && ng == G && ne == E) {                                                                       \
            done = true;                                                                                         \
            if (L == 16)                                                                                         \
                hipLaunchKernelGGL((k_leapfrog3d_spin_seg<G, E, 16>), dim3(grid), dim3(64), lds_s, s, p->dev, idx, A, \
                                   r0, v0, B_gradient, dt, n_steps, j0, j1, r_escape, r_traj, v_traj, r_final,   \
                                   v_final, n_actual);                                                           \
            else if (L == 4)                                                                                     \
                hipLaunchKernelGGL((k_leapfrog3d_spin_seg<G, E, 4>), dim3(grid), dim3(64), lds_s, s, p->dev, idx, A, \
                                   r0, v0, B_gradient, dt, n_steps, j0, j1, r_escape, r_traj, v_traj, r_final,   \
                                   v_final, n_actual);                                                           \
            else                                                                                                 \
                hipLaunchKernelGGL((k_leapfrog3d_spin_seg<G, E, 1>), dim3(grid), dim3(64), lds_s, s, p->dev, idx, A, \
                                   r0, v0, B_gradient, dt, n_steps, j0, j1, r_escape, r_traj, v_traj, r_final,   \
                                   v_final, n_actual);                                                           \
        }
        S3_SIZE_LIST(X)
#undef X
        if (!done) return mm_fail(MM_EUNSUPPORTED, "force3d_leapfrog_ws: no kernel");
        const int rc = mm_check_launch("k_leapfrog3d_spin_seg");
        if (rc != MM_OK) return rc;
        if (j1 >= n_steps) break;
        hipLaunchKernelGGL(k_s3_compact, dim3(1), dim3(1024), 0, s, idx, A, (const int*)n_actual, n_steps, out, cnt);
        int h = 0;
        if (hipMemcpyAsync(&h, cnt, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return mm_hip_fail("force3d_leapfrog_ws: alive count");
        A = h;
        idx = out;
        out = (out == ia) ? ib : ia;
    }
    return mm_check_launch("k_leapfrog3d_spin_seg");
}

}  // extern "C"
