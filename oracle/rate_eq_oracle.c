/*
 * CPU FP64 restatement of molmot's rate-equation force and 1D trajectory
 * integrator.  TEST INFRASTRUCTURE ONLY: this file is the parity checker and
 * the CPU baseline.  Only tests/, the smoke check and bench.py's CPU-baseline
 * leg load it; the product path (molmot_amd) never calls it.
 *
 * Two flavours of the reference algorithm are restated, because they differ
 * at the 1e-8 level through hbar and at the singular-matrix corner cases:
 *
 *   ORACLE_JIT   -- molmot/obe/rate_equations_jit.py:18-145 (_rate_eq_core):
 *                   one pass over beams (saturation applied afterwards,
 *                   :58-83), hand-written Gaussian elimination with partial
 *                   pivoting and the 1e-30 pivot-skip rule (:99-126),
 *                   back substitution leaving p[i]=0 on a tiny pivot
 *                   (:120-126), clamp and normalise only if the sum is
 *                   positive (:128-136).
 *   ORACLE_NUMPY -- molmot/obe/rate_equations.py:23-191 (solve_rate_equations):
 *                   two passes (s_total first, :103-127; saturated rates,
 *                   :129-154), per-beam rates summed over the beam axis
 *                   (:157), LAPACK-style LU solve (scipy.linalg.solve, :173)
 *                   with an exactly singular matrix mapped to the uniform
 *                   distribution (:174-175), clamp/normalise with the
 *                   all-zero case mapped to uniform (:177-182).
 *
 * hbar is an argument: the JIT path hard-codes 1.0545718e-34
 * (rate_equations_jit.py:27), the NumPy path uses CODATA 1.054571817e-34
 * (molmot/constants.py:18).
 *
 * `saturate` = 0 drops the 1/(1+s_total) factor, which is the unsaturated
 * model of validation/sroh_redmot_three_way.py:110-173 (the Julia algorithm);
 * it exists so the Julia fixture pins the Lorentzian / matrix / solve stages.
 *
 * Floating-point evaluation order follows the Python source expression by
 * expression; build with -ffp-contract=off so no FMA is formed.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_JIT 0
#define ORACLE_NUMPY 1
#define MAXG 64
#define MAXE 16
#define MAXB 64

static const double TWO_PI = 2.0 * 3.141592653589793;

/* numpy pairwise_sum for a contiguous run of n <= 128 doubles (the reduction
 * np.sum(R_exc, axis=2) at rate_equations.py:157 uses it). */
static double np_pairwise_sum(const double* a, int n) {
    if (n < 8) {
        double res = 0.0;
        for (int i = 0; i < n; i++) res += a[i];
        return res;
    }
    double r[8];
    for (int j = 0; j < 8; j++) r[j] = a[j];
    int i;
    for (i = 8; i < n - (n % 8); i += 8)
        for (int j = 0; j < 8; j++) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res += a[i];
    return res;
}

/* Branching ratios BR[ie][ig] (rate_equations_jit.py:37-49 and
 * rate_equations.py:84-93; the two differ only in summation grouping). */
static void branching_ratios(int n_g, int n_e, const double* d2, int flavour,
                             double BR[MAXE][MAXG]) {
    for (int ie = 0; ie < n_e; ie++) {
        double total = 0.0;
        for (int ig = 0; ig < n_g; ig++) {
            const double* d = d2 + (ig * n_e + ie) * 3;
            if (flavour == ORACLE_JIT) {
                for (int q = 0; q < 3; q++) total += d[q];
            } else {
                total += np_pairwise_sum(d, 3);
            }
        }
        for (int ig = 0; ig < n_g; ig++) {
            BR[ie][ig] = 0.0;
        }
        if (total > 1e-30) {
            for (int ig = 0; ig < n_g; ig++) {
                const double* d = d2 + (ig * n_e + ie) * 3;
                double s;
                if (flavour == ORACLE_JIT) {
                    s = 0.0;
                    for (int q = 0; q < 3; q++) s += d[q];
                } else {
                    s = np_pairwise_sum(d, 3);
                }
                BR[ie][ig] = s / total;
            }
        }
    }
}

/* Gaussian elimination with partial pivoting and the 1e-30 skip rule,
 * rate_equations_jit.py:99-126.  A is n x n row-major, overwritten. */
static void ge_jit(int n, double* A, double* b, double* p) {
    for (int col = 0; col < n; col++) {
        double max_val = fabs(A[col * n + col]);
        int max_row = col;
        for (int row = col + 1; row < n; row++) {
            if (fabs(A[row * n + col]) > max_val) {
                max_val = fabs(A[row * n + col]);
                max_row = row;
            }
        }
        if (max_row != col) {
            for (int j = 0; j < n; j++) {
                double t = A[col * n + j];
                A[col * n + j] = A[max_row * n + j];
                A[max_row * n + j] = t;
            }
            double t = b[col]; b[col] = b[max_row]; b[max_row] = t;
        }
        for (int row = col + 1; row < n; row++) {
            if (fabs(A[col * n + col]) < 1e-30) continue;
            double factor = A[row * n + col] / A[col * n + col];
            for (int j = col; j < n; j++) A[row * n + j] -= factor * A[col * n + j];
            b[row] -= factor * b[col];
        }
    }
    for (int i = 0; i < n; i++) p[i] = 0.0;
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int j = i + 1; j < n; j++) s -= A[i * n + j] * p[j];
        if (fabs(A[i * n + i]) > 1e-30) p[i] = s / A[i * n + i];
    }
}

/* LU with partial pivoting as LAPACK dgetf2 + dgetrs (scipy.linalg.solve at
 * rate_equations.py:173): first-maximum pivot, multipliers scaled by the
 * reciprocal of the pivot, rank-1 update, then L and U triangular solves.
 * Returns 0, or 1 when a pivot is exactly zero (LinAlgError). */
static int lu_solve_lapack(int n, double* A, double* b) {
    int piv[MAXG];
    for (int j = 0; j < n; j++) {
        int jp = j;
        double amax = fabs(A[j * n + j]);
        for (int i = j + 1; i < n; i++) {
            if (fabs(A[i * n + j]) > amax) { amax = fabs(A[i * n + j]); jp = i; }
        }
        piv[j] = jp;
        if (A[jp * n + j] == 0.0) return 1;
        if (jp != j) {
            for (int c = 0; c < n; c++) {
                double t = A[j * n + c]; A[j * n + c] = A[jp * n + c]; A[jp * n + c] = t;
            }
        }
        double r = 1.0 / A[j * n + j];
        for (int i = j + 1; i < n; i++) A[i * n + j] *= r;
        for (int i = j + 1; i < n; i++) {
            double l = A[i * n + j];
            for (int c = j + 1; c < n; c++) A[i * n + c] -= l * A[j * n + c];
        }
    }
    for (int j = 0; j < n; j++) {
        int jp = piv[j];
        if (jp != j) { double t = b[j]; b[j] = b[jp]; b[jp] = t; }
    }
    for (int i = 0; i < n; i++) {
        double s = b[i];
        for (int c = 0; c < i; c++) s -= A[i * n + c] * b[c];
        b[i] = s;
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int c = i + 1; c < n; c++) s -= A[i * n + c] * b[c];
        b[i] = s / A[i * n + i];
    }
    return 0;
}

/*
 * One rate-equation evaluation at (v, z).
 * d_squared: (n_g, n_e, 3) row-major; beam_pols: 0,1,2 (q = -1, 0, +1).
 * Outputs F (N), p (n_g), R (rad/s).  Returns 0, or -1 on bad sizes.
 */
int oracle_rate_eq(int n_g, int n_e, int n_b,
                   const double* energies, const double* d_squared,
                   const double* zeeman_z_diag,
                   const double* beam_dirs, const double* beam_freqs,
                   const int* beam_pols, const double* beam_s0s,
                   double v, double z, double B_gradient, double Gamma,
                   double k_wave, double Gamma_eff, double hbar,
                   int flavour, int saturate,
                   double* F_out, double* p_out, double* R_out) {
    if (n_g < 1 || n_g > MAXG || n_e < 1 || n_e > MAXE || n_b < 0 || n_b > MAXB) return -1;
    const int n_s = n_g + n_e;
    double Es[MAXG + MAXE];
    double BR[MAXE][MAXG];

    /* Zeeman-shifted energies: rate_equations_jit.py:29-35 evaluates
     * E + ((B*zd)*Gamma)/(2 pi); rate_equations.py:76-82 evaluates
     * E + ((B*zd)*Gamma)/(2 pi) as an array expression -- same order. */
    double B_gauss;
    if (flavour == ORACLE_JIT) {
        B_gauss = B_gradient * z * 1e4;
    } else {
        double B_tesla = B_gradient * z;
        B_gauss = B_tesla * 1e4;
    }
    for (int i = 0; i < n_s; i++)
        Es[i] = energies[i] + B_gauss * zeeman_z_diag[i] * Gamma / TWO_PI;

    branching_ratios(n_g, n_e, d_squared, flavour, BR);

    const double gamma_half = Gamma / 2.0;
    /* (Gamma/2)**2 and delta**2 in the NumPy path are correctly rounded
     * squares, identical to x*x */
    const double gamma_half_sq = gamma_half * gamma_half;
    double R_sum[MAXG][MAXE];
    double F_tot[MAXG][MAXE];
    double s_total[MAXG];
    for (int ig = 0; ig < n_g; ig++) {
        s_total[ig] = 0.0;
        for (int ie = 0; ie < n_e; ie++) R_sum[ig][ie] = F_tot[ig][ie] = 0.0;
    }

    if (flavour == ORACLE_JIT) {
        for (int ib = 0; ib < n_b; ib++) {
            double kdir = beam_dirs[ib], freq = beam_freqs[ib], s0 = beam_s0s[ib];
            int pol = beam_pols[ib];
            double doppler = -k_wave * kdir * v / TWO_PI;
            for (int ig = 0; ig < n_g; ig++) {
                for (int ie = 0; ie < n_e; ie++) {
                    double d2 = d_squared[(ig * n_e + ie) * 3 + pol];
                    if (d2 < 1e-15) continue;
                    double omega_trans = Es[n_g + ie] - Es[ig];
                    double delta_eff = (freq + doppler - omega_trans) * TWO_PI;
                    double L = gamma_half_sq / (delta_eff * delta_eff + gamma_half_sq);
                    double rate_unsat = gamma_half * s0 * d2 * L * Gamma_eff;
                    R_sum[ig][ie] += rate_unsat;
                    F_tot[ig][ie] += hbar * k_wave * kdir * rate_unsat;
                    s_total[ig] += s0 * d2 * L;
                }
            }
        }
        if (saturate) {
            for (int ig = 0; ig < n_g; ig++) {
                double sat_corr = 1.0 / (1.0 + s_total[ig]);
                for (int ie = 0; ie < n_e; ie++) {
                    R_sum[ig][ie] *= sat_corr;
                    F_tot[ig][ie] *= sat_corr;
                }
            }
        }
    } else {
        /* pass 1: s_total (rate_equations.py:107-127) */
        if (saturate) {
            for (int ib = 0; ib < n_b; ib++) {
                double kdir = beam_dirs[ib];
                double doppler = -k_wave * kdir * v / TWO_PI;
                int q = beam_pols[ib];
                for (int ig = 0; ig < n_g; ig++) {
                    for (int ie = 0; ie < n_e; ie++) {
                        double d2 = d_squared[(ig * n_e + ie) * 3 + q];
                        if (d2 < 1e-15) continue;
                        double omega_trans = Es[n_g + ie] - Es[ig];
                        double delta_eff = (beam_freqs[ib] + doppler - omega_trans) * TWO_PI;
                        double L = gamma_half_sq / (delta_eff * delta_eff + gamma_half_sq);
                        s_total[ig] += beam_s0s[ib] * d2 * L;
                    }
                }
            }
        }
        /* pass 2: per-beam rates, then summed over the beam axis (:130-157) */
        double* R_exc = (double*)calloc((size_t)n_g * n_e * (n_b > 0 ? n_b : 1), sizeof(double));
        double* F_per = (double*)calloc((size_t)n_g * n_e * (n_b > 0 ? n_b : 1), sizeof(double));
        for (int ib = 0; ib < n_b; ib++) {
            double kdir = beam_dirs[ib];
            double doppler = -k_wave * kdir * v / TWO_PI;
            int q = beam_pols[ib];
            for (int ig = 0; ig < n_g; ig++) {
                for (int ie = 0; ie < n_e; ie++) {
                    double d2 = d_squared[(ig * n_e + ie) * 3 + q];
                    if (d2 < 1e-15) continue;
                    double omega_trans = Es[n_g + ie] - Es[ig];
                    double delta_eff = (beam_freqs[ib] + doppler - omega_trans) * TWO_PI;
                    double L = gamma_half_sq / (delta_eff * delta_eff + gamma_half_sq);
                    double rate = gamma_half * beam_s0s[ib] * d2 * L * Gamma_eff;
                    if (saturate) rate = rate / (1.0 + s_total[ig]);
                    R_exc[(ig * n_e + ie) * n_b + ib] = rate;
                    F_per[(ig * n_e + ie) * n_b + ib] = hbar * k_wave * kdir * rate;
                }
            }
        }
        for (int ig = 0; ig < n_g; ig++)
            for (int ie = 0; ie < n_e; ie++) {
                R_sum[ig][ie] = np_pairwise_sum(R_exc + (ig * n_e + ie) * n_b, n_b);
            }
        /* force = sum_ig p[ig] * np.sum(F_per[ig,:,:]) (:187-189): keep the
         * per-ground total of F_per, reduced over the flattened (n_e*n_b) axis */
        for (int ig = 0; ig < n_g; ig++) {
            F_tot[ig][0] = np_pairwise_sum(F_per + ig * n_e * n_b, n_e * n_b);
            for (int ie = 1; ie < n_e; ie++) F_tot[ig][ie] = 0.0;
        }
        free(R_exc);
        free(F_per);
    }

    /* population matrix (rate_equations_jit.py:85-97; rate_equations.py:159-170) */
    const int n = n_g;
    double M[MAXG * MAXG];
    double b[MAXG], p[MAXG];
    for (int i = 0; i < n * n; i++) M[i] = 0.0;
    for (int ig = 0; ig < n_g; ig++) {
        if (flavour == ORACLE_JIT) {
            for (int ie = 0; ie < n_e; ie++) M[ig * n + ig] -= R_sum[ig][ie];
        } else {
            M[ig * n + ig] -= np_pairwise_sum(R_sum[ig], n_e);
        }
        for (int ik = 0; ik < n_g; ik++)
            for (int ie = 0; ie < n_e; ie++)
                M[ig * n + ik] += R_sum[ik][ie] * BR[ie][ig];
    }
    for (int j = 0; j < n; j++) M[(n - 1) * n + j] = 1.0;
    for (int i = 0; i < n; i++) b[i] = 0.0;
    b[n - 1] = 1.0;

    if (flavour == ORACLE_JIT) {
        ge_jit(n, M, b, p);
        for (int i = 0; i < n; i++) if (p[i] < 0) p[i] = 0.0;
        double total_p = 0.0;
        for (int i = 0; i < n; i++) total_p += p[i];
        if (total_p > 0) for (int i = 0; i < n; i++) p[i] /= total_p;
    } else {
        int singular = lu_solve_lapack(n, M, b);
        if (singular) {
            for (int i = 0; i < n; i++) p[i] = 1.0 / n;
        } else {
            for (int i = 0; i < n; i++) p[i] = b[i];
        }
        for (int i = 0; i < n; i++) p[i] = p[i] > 0.0 ? p[i] : 0.0;
        double p_sum = np_pairwise_sum(p, n);
        if (p_sum > 0) {
            for (int i = 0; i < n; i++) p[i] /= p_sum;
        } else {
            for (int i = 0; i < n; i++) p[i] = 1.0 / n;
        }
    }

    double force = 0.0, R_scatter = 0.0;
    for (int ig = 0; ig < n_g; ig++) {
        if (flavour == ORACLE_JIT) {
            for (int ie = 0; ie < n_e; ie++) {
                force += p[ig] * F_tot[ig][ie];
                R_scatter += p[ig] * R_sum[ig][ie];
            }
        } else {
            force += p[ig] * F_tot[ig][0];
            R_scatter += p[ig] * np_pairwise_sum(R_sum[ig], n_e);
        }
    }
    *F_out = force;
    *R_out = R_scatter;
    if (p_out) for (int i = 0; i < n; i++) p_out[i] = p[i];
    return 0;
}

/* Batched evaluation over n scattered (v[i], z[i]) points, OpenMP-parallel.
 * P may be NULL; otherwise it is (n, n_g) row-major. */
int oracle_rate_eq_batch(int n_g, int n_e, int n_b,
                         const double* energies, const double* d_squared,
                         const double* zeeman_z_diag,
                         const double* beam_dirs, const double* beam_freqs,
                         const int* beam_pols, const double* beam_s0s,
                         long n, const double* v, const double* z,
                         double B_gradient, double Gamma, double k_wave,
                         double Gamma_eff, double hbar, int flavour, int saturate,
                         double* F, double* P, double* R, int n_threads) {
    int err = 0;
    if (n_threads < 1) n_threads = 1;
#pragma omp parallel for schedule(dynamic, 64) num_threads(n_threads) reduction(|:err)
    for (long i = 0; i < n; i++) {
        err |= oracle_rate_eq(n_g, n_e, n_b, energies, d_squared, zeeman_z_diag,
                              beam_dirs, beam_freqs, beam_pols, beam_s0s,
                              v[i], z[i], B_gradient, Gamma, k_wave, Gamma_eff, hbar,
                              flavour, saturate, &F[i], P ? P + i * n_g : NULL, &R[i]);
    }
    return err;
}

/*
 * 1D symplectic-Euler trajectories, molmot/propagation/trajectories_jit.py:18-58
 * (_trajectory_1d_jit) with the force from the JIT-flavour evaluation:
 *   v_j = v_{j-1} + (F(v_{j-1}, z_{j-1}) / m) dt;  z_j = z_{j-1} + v_j dt
 * for j = 1 .. n_steps-1, stopping at the first |z_j| > z_escape
 * (n_actual = j; the remaining samples repeat the escape state, :51-56).
 * z_traj / v_traj may be NULL; otherwise (n, n_steps) row-major.
 */
int oracle_traj_euler(int n_g, int n_e, int n_b,
                      const double* energies, const double* d_squared,
                      const double* zeeman_z_diag,
                      const double* beam_dirs, const double* beam_freqs,
                      const int* beam_pols, const double* beam_s0s,
                      long n, const double* z0, const double* v0,
                      double dt, int n_steps, double mass,
                      double B_gradient, double Gamma, double k_wave,
                      double Gamma_eff, double hbar, double z_escape, int flavour,
                      double* zf, double* vf, int* n_actual,
                      double* z_traj, double* v_traj, int n_threads) {
    int err = 0;
    if (n_threads < 1) n_threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads) reduction(|:err)
    for (long i = 0; i < n; i++) {
        double z = z0[i], v = v0[i];
        int na = n_steps;
        if (z_traj) { z_traj[i * (long)n_steps] = z; v_traj[i * (long)n_steps] = v; }
        for (int j = 1; j < n_steps; j++) {
            double F, R;
            err |= oracle_rate_eq(n_g, n_e, n_b, energies, d_squared, zeeman_z_diag,
                                  beam_dirs, beam_freqs, beam_pols, beam_s0s,
                                  v, z, B_gradient, Gamma, k_wave, Gamma_eff, hbar,
                                  flavour, 1, &F, NULL, &R);
            double a = F / mass;
            v = v + a * dt;
            z = z + v * dt;
            if (z_traj) { z_traj[i * (long)n_steps + j] = z; v_traj[i * (long)n_steps + j] = v; }
            if (fabs(z) > z_escape) {
                na = j;
                if (z_traj) {
                    for (int k = j + 1; k < n_steps; k++) {
                        z_traj[i * (long)n_steps + k] = z;
                        v_traj[i * (long)n_steps + k] = v;
                    }
                }
                break;
            }
        }
        zf[i] = z;
        vf[i] = v;
        n_actual[i] = na;
    }
    return err;
}

/*
 * 3D per-axis trajectories, molmot/propagation/trajectories.py:78-156
 * (simulate_trajectories_3d): each step evaluates the 1D force three times,
 *   Fz = F(v_z, z), Fx = F(v_x, -x/2), Fy = F(v_y, -y/2)       (:125-145)
 * then v += F/m dt, r += v dt, and stops at the first |r| > r_escape
 * (n_actual = j; saved samples after it repeat the escape state, :150-153).
 * r0, v0, r_final, v_final: (n, 3); r_traj / v_traj: (n, n_steps, 3) or NULL.
 */
int oracle_traj_axis3d(int n_g, int n_e, int n_b,
                       const double* energies, const double* d_squared,
                       const double* zeeman_z_diag,
                       const double* beam_dirs, const double* beam_freqs,
                       const int* beam_pols, const double* beam_s0s,
                       long n, const double* r0, const double* v0,
                       double dt, int n_steps, double mass,
                       double B_gradient, double Gamma, double k_wave,
                       double Gamma_eff, double hbar, double r_escape, int flavour,
                       double* rf, double* vf, int* n_actual,
                       double* r_traj, double* v_traj, int n_threads) {
    int err = 0;
    if (n_threads < 1) n_threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads) reduction(|:err)
    for (long i = 0; i < n; i++) {
        double r[3] = {r0[3 * i], r0[3 * i + 1], r0[3 * i + 2]};
        double v[3] = {v0[3 * i], v0[3 * i + 1], v0[3 * i + 2]};
        int na = n_steps;
        long base = i * (long)n_steps * 3;
        if (r_traj) for (int a = 0; a < 3; a++) { r_traj[base + a] = r[a]; v_traj[base + a] = v[a]; }
        for (int j = 1; j < n_steps; j++) {
            double F[3], R;
            err |= oracle_rate_eq(n_g, n_e, n_b, energies, d_squared, zeeman_z_diag, beam_dirs, beam_freqs,
                                  beam_pols, beam_s0s, v[2], r[2], B_gradient, Gamma, k_wave, Gamma_eff, hbar,
                                  flavour, 1, &F[2], NULL, &R);
            err |= oracle_rate_eq(n_g, n_e, n_b, energies, d_squared, zeeman_z_diag, beam_dirs, beam_freqs,
                                  beam_pols, beam_s0s, v[0], -r[0] / 2.0, B_gradient, Gamma, k_wave, Gamma_eff,
                                  hbar, flavour, 1, &F[0], NULL, &R);
            err |= oracle_rate_eq(n_g, n_e, n_b, energies, d_squared, zeeman_z_diag, beam_dirs, beam_freqs,
                                  beam_pols, beam_s0s, v[1], -r[1] / 2.0, B_gradient, Gamma, k_wave, Gamma_eff,
                                  hbar, flavour, 1, &F[1], NULL, &R);
            for (int a = 0; a < 3; a++) {
                double acc = F[a] / mass;
                v[a] = v[a] + acc * dt;
                r[a] = r[a] + v[a] * dt;
            }
            if (r_traj) for (int a = 0; a < 3; a++) {
                r_traj[base + (long)j * 3 + a] = r[a];
                v_traj[base + (long)j * 3 + a] = v[a];
            }
            if (sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) > r_escape) {
                na = j;
                if (r_traj)
                    for (int k = j + 1; k < n_steps; k++)
                        for (int a = 0; a < 3; a++) {
                            r_traj[base + (long)k * 3 + a] = r[a];
                            v_traj[base + (long)k * 3 + a] = v[a];
                        }
                break;
            }
        }
        for (int a = 0; a < 3; a++) { rf[3 * i + a] = r[a]; vf[3 * i + a] = v[a]; }
        n_actual[i] = na;
    }
    return err;
}

/*
 * Monte-Carlo temperature statistics, molmot/propagation/trajectories.py:159-229:
 * Euler steps j = 1 .. n_steps-1; a particle stops after the first step with
 * |z| > 0.015 (z_escape); at every other step with j >= i_eq, v^2 and one
 * sample are accumulated.  Per-particle sums are returned (the reference
 * accumulates one running sum over all particles; T = m * sum / (n k_B)).
 */
int oracle_traj_vsq(int n_g, int n_e, int n_b,
                    const double* energies, const double* d_squared,
                    const double* zeeman_z_diag,
                    const double* beam_dirs, const double* beam_freqs,
                    const int* beam_pols, const double* beam_s0s,
                    long n, const double* z0, const double* v0,
                    double dt, int n_steps, int i_eq, double mass,
                    double B_gradient, double Gamma, double k_wave,
                    double Gamma_eff, double hbar, double z_escape, int flavour,
                    double* vsq_sum, long* n_samples, int n_threads) {
    int err = 0;
    if (n_threads < 1) n_threads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads) reduction(|:err)
    for (long i = 0; i < n; i++) {
        double z = z0[i], v = v0[i], acc = 0.0;
        long cnt = 0;
        for (int j = 1; j < n_steps; j++) {
            double F, R;
            err |= oracle_rate_eq(n_g, n_e, n_b, energies, d_squared, zeeman_z_diag, beam_dirs, beam_freqs,
                                  beam_pols, beam_s0s, v, z, B_gradient, Gamma, k_wave, Gamma_eff, hbar,
                                  flavour, 1, &F, NULL, &R);
            double a = F / mass;
            v += a * dt;
            z += v * dt;
            if (fabs(z) > z_escape) break;
            if (j >= i_eq) {
                acc += v * v;
                cnt += 1;
            }
        }
        vsq_sum[i] = acc;
        n_samples[i] = cnt;
    }
    return err;
}

/* Test utility for the kernels' division (x / y from RN(1/y) with one FMA
 * correction, Markstein): over n pseudo-random x (xorshift64 from seed, random
 * sign and significand, exponents -60..80) count the x for which
 * fma(x - q0 y, yi, q0), q0 = RN(x yi), yi = RN(1/y), differs from the IEEE
 * quotient x / y.  Zero is expected for every y away from overflow and the
 * subnormal range.  correct = 0 drops the correction step (the plain product
 * x * RN(1/y)), a planted mistake the count must expose. */
long oracle_div_rn_mismatches(double y, long n, unsigned long long seed, int correct) {
    volatile double yv = y;
    const double yi = 1.0 / yv;
    unsigned long long s = seed ? seed : 88172645463325252ULL;
    long bad = 0;
    for (long i = 0; i < n; i++) {
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        const unsigned long long m = s & ((1ULL << 52) - 1);
        s ^= s << 13;
        s ^= s >> 7;
        s ^= s << 17;
        const int e = -60 + (int)(s % 141ULL);
        const unsigned long long bits = ((unsigned long long)(e + 1023) << 52) | m | ((s >> 40 & 1ULL) << 63);
        double x;
        memcpy(&x, &bits, sizeof(x));
        const double q0 = x * yi;
        const double r = fma(-q0, yv, x);
        const double q = correct ? fma(r, yi, q0) : q0;
        if (q != x / yv) bad++;
    }
    return bad;
}
