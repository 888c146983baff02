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
    const int ng = p->host.ng, ne = p->host.ne;
#define X(G, E)                                                                                              \
    if (ng == G && ne == E) {                                                                                \
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
    const int ng = p->host.ng, ne = p->host.ne;
#define X(G, E)                                                                                                 \
    if (ng == G && ne == E) {                                                                                   \
        hipLaunchKernelGGL((k_leapfrog3d<G, E>), dim3(grid), dim3(64), 0, (hipStream_t)stream, p->dev, n, r0, v0, \
                           B_gradient, dt, n_steps, r_escape, r_traj, v_traj, r_final, v_final, n_actual);       \
        return mm_check_launch("k_leapfrog3d");                                                                 \
    }
    S3_SIZE_LIST(X)
#undef X
    return mm_fail(MM_EUNSUPPORTED, "force3d_leapfrog: no kernel");
}

}  // extern "C"
