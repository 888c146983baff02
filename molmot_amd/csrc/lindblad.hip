// gfx950 kernels and C ABI for the batched standing-wave Lindblad propagator.
//
// Model (natural units: t in 1/Gamma, H in Gamma), following
// molmot/obe/obe_subdoppler.py:19-141:
//   H(t) = diag(Hd) + C(t) + C(t)^dagger,  C(t) = e^{i phi} A + e^{-i phi} B
//   A, B: forward / backward couplings, nonzero in the (excited, ground) block
//   phi(t) = phi0 + kv t   (kv = k v / Gamma: the moving standing-wave phase;
//                           the reference's Hamiltonian ignores v, :85-97)
//   d rho/dt = -i[H, rho] - (gam_i + gam_j)/2 rho_ij
//              + delta_ij [i ground] sum_e rep[i][e] rho_ee
// the dissipator of the reference's jump operators sqrt(Gamma_eff br)|g><e|
// (:124-139).  The force sample is (:186-228)
//   F = 2 hbar k Gamma Im sum_{g,e} rho_ge (e^{i phi} A_eg - e^{-i phi} B_eg).
// Classic RK4 from rho(0) (all ground states equally populated); F and the
// excited population are averaged over the step ends after `avg_start`,
// accumulated per lane in FP64 (Neumaier-compensated in the VALU back end).
//
// Two contraction back ends:
//   k_lindblad_rk4_valu<T, NG, NE>: X = H rho with the block structure of H
//     (a ground row needs NE terms, an excited row NG terms) on VALU, in
//     T = double (the FP64 path) or float, 64 to 256 lanes per instance;
//   k_lindblad_rk4_mfma16: 12 + 4 levels, T = float, one wavefront per
//     instance: X = H rho on v_mfma_f32_16x16x4f32 in two parts, the
//     excited rows (K = 12 ground rows, real and imaginary parts stacked in
//     one M tile) and the ground rows (K = 4 excited rows): 10 MFMA per
//     product.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include "../../include/molmot_hip.h"
#include "mm_errors.h"
#include "mm_trig.h"

#define LB_MAXG 36
#define LB_MAXE 4

struct LbClass {
    int ng, ne;
    double Hd[LB_MAXG + LB_MAXE];
    double Ar[LB_MAXE][LB_MAXG], Ai[LB_MAXE][LB_MAXG];  // A[e][g]
    double Br[LB_MAXE][LB_MAXG], Bi[LB_MAXE][LB_MAXG];  // B[e][g]
    double gam[LB_MAXE];                                // decay of excited e
    double rep[LB_MAXG][LB_MAXE];                       // e -> g repopulation
    double force_unit;                                  // 2 hbar k Gamma (N)
    double rate_unit;                                   // Gamma (1/s)
    // force weights of rho_ge = x + i y with (c, s) = (cos, sin) of the phase:
    //   Im[rho_ge (e^{i phi} A_eg - e^{-i phi} B_eg)] = c (x Q2 + y P1) + s (x P2 - y Q1)
    //   fw[e][g] = {P1, Q1, P2, Q2} = {Ar - Br, Ai + Bi, Ar + Br, Ai - Bi}
    double fw[LB_MAXE][LB_MAXG][4];
};

struct mm_lindblad_class {
    LbClass host;
    LbClass* dev;
    int lanes[2];  // lanes per instance of the VALU back end, FP64 / FP32 (0: default)
};

__device__ __forceinline__ void neumaier(double& s, double& c, double x) {
    const double t = s + x;
    c += (fabs(s) >= fabs(x)) ? (s - t) + x : (x - t) + s;
    s = t;
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Hides the class pointer from loop-invariant code motion: per-lane table
// entries (diagonal energies, decay and repopulation rates, force weights)
// are then loaded where they are used instead of being hoisted out of the
// step loop and held in registers for every owned element.
__device__ __forceinline__ const LbClass* lb_fresh(const LbClass* p) {
    asm volatile("" : "+s"(p));
    return p;
}

// The same for the lane index: the element indices and LDS addresses of the
// owned elements are recomputed per step rather than held for the whole run.
__device__ __forceinline__ int lb_fresh_lane(int lane) {
    asm volatile("" : "+v"(lane));
    return lane;
}

// A wave-uniform double (computed by the vector unit) moved to scalar
// registers: read by every lane, it then costs no vector registers.
__device__ __forceinline__ double lb_uniform(double x) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readfirstlane((int)b);
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// a*b + c with one rounding.  The library is compiled with contraction off
// (the rate-equation detunings need the reference's roundings), so the
// propagators write their fused products out.
__device__ __forceinline__ float lb_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double lb_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// RK4 state update x += h k.  FP64: one fused multiply-add.  FP32: the
// increment carries the previous step's rounding error `lo` and the new one
// is kept by Knuth's two-sum (x + lo is the state to ~FP32^2), so rounding
// does not accumulate over the ~1e4 steps of a run along the slow modes
// (ground-state coherences and populations, which relax only by optical
// pumping).
__device__ __forceinline__ void lb_update(double& x, double& lo, double h, double k) {
    (void)lo;
    x = __builtin_fma(h, k, x);
}
__device__ __forceinline__ void lb_update(float& x, float& lo, float h, float k) {
    const float inc = __builtin_fmaf(h, k, lo);
    const float s = x + inc;
    const float bb = s - x;
    lo = (x - (s - bb)) + (inc - bb);
    x = s;
}


// Sum over the BS lanes of one instance: wavefront butterfly, then the
// per-wave partials through LDS (red holds BS / 64 doubles per quantity).
template <int BS>
__device__ __forceinline__ void lb_group_sum2(double& a, double& b, int lane, double* red) {
    a = wave_sum(a);
    b = wave_sum(b);
    if (BS > 64) {
        constexpr int W = BS / 64;
        if ((lane & 63) == 0) {
            red[lane >> 6] = a;
            red[W + (lane >> 6)] = b;
        }
        __syncthreads();
        a = b = 0.0;
#pragma unroll
        for (int w = 0; w < W; w++) {
            a += red[w];
            b += red[W + w];
        }
    }
}

// ---------------------------------------------------------------------------
// VALU back end: one instance per group of BS lanes (BS = 64, 128, 256: the
// lanes-per-instance choice; larger groups hold fewer elements per lane).
// lds: 4 N^2 + 2 NE NG elements of T; red: 2 BS / 64 doubles.
// `active` false runs the instance (so every barrier is met) but writes nothing.
// ---------------------------------------------------------------------------
template <typename T, int NG, int NE, int BS>
__device__ __forceinline__ void lb_valu_body(const LbClass* __restrict__ C, long inst, bool active, int lane0,
                                             double kvi, double ph0, double dt, int n_steps, int avg_start,
                                             double* __restrict__ F_out, double* __restrict__ R_out,
                                             T* __restrict__ rho_out, T* __restrict__ lds, double* red) {
    constexpr int N = NG + NE;
    constexpr int NN = N * N;
    constexpr int M = (NN + BS - 1) / BS;  // owned elements per lane
    T* const sr = lds;
    T* const si = lds + NN;
    // X = C rho is read back transposed: a row pitch of N + 1 elements puts
    // the column a wavefront reads on distinct LDS banks (pitch N put every
    // lane of a 16-element column on one bank: 46% of the FP64 kernel's LDS
    // cycles were bank conflicts)
    constexpr int XP = N + 1;
    T* const xr = lds + 2 * NN;
    T* const xi = xr + N * XP;
    T* const hr = xi + N * XP;
    T* const hi = hr + NE * NG;

    int lane = lane0;
    // rl, il: the rounding error of the FP32 state (compensated RK4 update;
    // unused in FP64)
    T rr[M], ri[M], ar[M], ai[M], rl[M], il[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int idx = lane + BS * m;
        const int i = idx / N, j = idx - (idx / N) * N;
        rr[m] = (idx < NN && i == j && i < NG) ? (T)(1.0 / NG) : (T)0;
        ri[m] = (T)0;
        rl[m] = il[m] = (T)0;
    }

    // f(stage in LDS) -> (kr, ki) for the owned elements; (c, s) = (cos, sin)
    // of the standing-wave phase at the stage time
    auto rhs = [&](const T c, const T s, T (&kr)[M], T (&ki)[M]) {
        const int lane = lb_fresh_lane(lane0);
        // small systems keep the per-lane table entries in registers; large
        // ones reload them per stage
        const LbClass* __restrict__ Cc = C;
        if constexpr (M > 4) Cc = lb_fresh(C);
        for (int q = lane; q < NE * NG; q += BS) {
            const int e = q / NG, g = q - (q / NG) * NG;
            const T a_r = (T)Cc->Ar[e][g], a_i = (T)Cc->Ai[e][g];
            const T b_r = (T)Cc->Br[e][g], b_i = (T)Cc->Bi[e][g];
            // e^{i ph} A + e^{-i ph} B
            hr[q] = lb_fma(c, a_r + b_r, s * (b_i - a_i));
            hi[q] = lb_fma(s, a_r - b_r, c * (a_i + b_i));
        }
        __syncthreads();
        // rolled for large M: unrolled, the compiler keeps the LDS coupling
        // rows of every owned element live at once
#pragma unroll(M <= 4 ? M : 1)
        for (int m = 0; m < M; m++) {
            const int idx = lane + BS * m;
            if (idx >= NN) break;
            const int i = idx / N, j = idx - (idx / N) * N;
            T x_r = (T)0, x_i = (T)0;   // (C rho)_ij: the diagonal of H enters below
            if (i < NG) {
#pragma unroll
                for (int e = 0; e < NE; e++) {  // H[i][NG+e] = conj(Hoff[e][i])
                    const T h_r = hr[e * NG + i], h_i = -hi[e * NG + i];
                    const T p_r = sr[(NG + e) * N + j], p_i = si[(NG + e) * N + j];
                    x_r = lb_fma(h_r, p_r, lb_fma(-h_i, p_i, x_r));
                    x_i = lb_fma(h_r, p_i, lb_fma(h_i, p_r, x_i));
                }
            } else {
                const int e = i - NG;
#pragma unroll 4
                for (int g = 0; g < NG; g++) {
                    const T h_r = hr[e * NG + g], h_i = hi[e * NG + g];
                    const T p_r = sr[g * N + j], p_i = si[g * N + j];
                    x_r = lb_fma(h_r, p_r, lb_fma(-h_i, p_i, x_r));
                    x_i = lb_fma(h_r, p_i, lb_fma(h_i, p_r, x_i));
                }
            }
            xr[i * XP + j] = x_r;
            xi[i * XP + j] = x_i;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int idx = lane + BS * m;
            if (idx >= NN) {
                kr[m] = ki[m] = (T)0;
                continue;
            }
            const int i = idx / N, j = idx - (idx / N) * N;
            const int xdx = i * XP + j, tdx = j * XP + i;
            // -i (X_ij - conj(X_ji)) - i (Hd_i - Hd_j) rho_ij; the energy
            // difference is formed in FP64 and rounded once: formed from two
            // FP32 products it would lose the slow ground-state coherences of
            // nearly degenerate levels to cancellation
            const T dh = (T)(Cc->Hd[i] - Cc->Hd[j]);
            T k_r = lb_fma(dh, si[idx], xi[xdx] + xi[tdx]);
            T k_i = lb_fma(-dh, sr[idx], xr[tdx] - xr[xdx]);
            const T gi = (i >= NG) ? (T)Cc->gam[i - NG] : (T)0;
            const T gj = (j >= NG) ? (T)Cc->gam[j - NG] : (T)0;
            const T dec = (T)0.5 * (gi + gj);
            k_r = lb_fma(-dec, sr[idx], k_r);
            k_i = lb_fma(-dec, si[idx], k_i);
            if (i == j && i < NG) {
#pragma unroll
                for (int e = 0; e < NE; e++) k_r = lb_fma((T)Cc->rep[i][e], sr[(NG + e) * N + NG + e], k_r);
            }
            kr[m] = k_r;
            ki[m] = k_i;
        }
        __syncthreads();  // the stage may be overwritten next
    };

    double fs = 0.0, fc = 0.0, rs = 0.0, rc = 0.0;
    const T h = (T)dt, h2 = (T)(0.5 * dt), h6 = (T)(dt / 6.0);
    // (cos, sin) of the standing-wave phase at the stage times: two
    // evaluations per step (t + dt/2 for stages 2 and 3, t + dt for stage 4,
    // the force sample and the next step's stage 1)
    double c0d, s0d;
    mm_cis(ph0, c0d, s0d);
    for (int step = 0; step < n_steps; step++) {
        const double t = step * dt;
        double chd, shd, c1d, s1d;
        mm_cis(ph0 + kvi * (t + 0.5 * dt), chd, shd);
        mm_cis(ph0 + kvi * (t + dt), c1d, s1d);
        const T c0 = (T)c0d, s0 = (T)s0d, ch = (T)chd, sh = (T)shd, c1 = (T)c1d, s1 = (T)s1d;
        lane = lb_fresh_lane(lane0);
        T kr[M], ki[M];
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int idx = lane + BS * m;
            if (idx < NN) {
                sr[idx] = rr[m];
                si[idx] = ri[m];
            }
            ar[m] = ai[m] = (T)0;
        }
        __syncthreads();
        rhs(c0, s0, kr, ki);  // k1
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int idx = lane + BS * m;
            ar[m] += kr[m];
            ai[m] += ki[m];
            if (idx < NN) {
                sr[idx] = lb_fma(h2, kr[m], rr[m]);
                si[idx] = lb_fma(h2, ki[m], ri[m]);
            }
        }
        __syncthreads();
        rhs(ch, sh, kr, ki);  // k2
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int idx = lane + BS * m;
            ar[m] = lb_fma((T)2, kr[m], ar[m]);
            ai[m] = lb_fma((T)2, ki[m], ai[m]);
            if (idx < NN) {
                sr[idx] = lb_fma(h2, kr[m], rr[m]);
                si[idx] = lb_fma(h2, ki[m], ri[m]);
            }
        }
        __syncthreads();
        rhs(ch, sh, kr, ki);  // k3
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int idx = lane + BS * m;
            ar[m] = lb_fma((T)2, kr[m], ar[m]);
            ai[m] = lb_fma((T)2, ki[m], ai[m]);
            if (idx < NN) {
                sr[idx] = lb_fma(h, kr[m], rr[m]);
                si[idx] = lb_fma(h, ki[m], ri[m]);
            }
        }
        __syncthreads();
        rhs(c1, s1, kr, ki);  // k4
#pragma unroll
        for (int m = 0; m < M; m++) lb_update(rr[m], rl[m], h6, ar[m] + kr[m]);
#pragma unroll
        for (int m = 0; m < M; m++) lb_update(ri[m], il[m], h6, ai[m] + ki[m]);
        if (step + 1 > avg_start) {
            const LbClass* __restrict__ Cs = lb_fresh(C);
            const double c = c1d, s = s1d;
            double f = 0.0, r = 0.0;
#pragma unroll
            for (int m = 0; m < M; m++) {
                const int idx = lane + BS * m;
                if (idx >= NN) break;
                const int i = idx / N, j = idx - (idx / N) * N;
                if (i < NG && j >= NG) {
                    const int e = j - NG;
                    // e^{i ph} A_eg - e^{-i ph} B_eg
                    const double ar_ = Cs->Ar[e][i], ai_ = Cs->Ai[e][i], br_ = Cs->Br[e][i], bi_ = Cs->Bi[e][i];
                    const double w_r = c * ar_ - s * ai_ - (c * br_ + s * bi_);
                    const double w_i = s * ar_ + c * ai_ - (c * bi_ - s * br_);
                    f += ((double)rr[m] + (double)rl[m]) * w_i + ((double)ri[m] + (double)il[m]) * w_r;
                } else if (i == j && i >= NG) {
                    r += (double)rr[m] + (double)rl[m];
                }
            }
            neumaier(fs, fc, f);
            neumaier(rs, rc, r);
        }
        c0d = c1d;
        s0d = s1d;
    }
    const double nsamp = (double)(n_steps - (avg_start < n_steps ? avg_start : n_steps));
    double F = fs + fc, R = rs + rc;
    lb_group_sum2<BS>(F, R, lane, red);
    if (!active) return;
    if (lane == 0) {
        F_out[inst] = nsamp > 0 ? C->force_unit * F / nsamp : 0.0;
        if (R_out) R_out[inst] = nsamp > 0 ? C->rate_unit * R / nsamp : 0.0;
    }
    if (rho_out) {
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int idx = lane + BS * m;
            if (idx < NN) {
                rho_out[(inst * NN + idx) * 2] = rr[m];
                rho_out[(inst * NN + idx) * 2 + 1] = ri[m];
            }
        }
    }
}

template <typename T, int NG, int NE>
constexpr int lb_valu_lds() {
    return 2 * (NG + NE) * (NG + NE) + 2 * (NG + NE) * (NG + NE + 1) + 2 * NE * NG;
}

template <typename T, int NG, int NE, int BS>
__global__ __launch_bounds__(BS) void k_lindblad_rk4_valu(const LbClass* __restrict__ C, long n_inst,
                                                          const double* __restrict__ kv,
                                                          const double* __restrict__ phi0, double dt, int n_steps,
                                                          int avg_start, double* __restrict__ F_out,
                                                          double* __restrict__ R_out, T* __restrict__ rho_out) {
    __shared__ T lds[lb_valu_lds<T, NG, NE>()];
    __shared__ double red[2 * (BS / 64)];
    const long inst = blockIdx.x;
    if (inst >= n_inst) return;  // uniform over the block
    lb_valu_body<T, NG, NE, BS>(C, inst, true, threadIdx.x, kv[inst], phi0[inst], dt, n_steps, avg_start, F_out,
                                R_out, rho_out, lds, red);
}


// ---------------------------------------------------------------------------
// MFMA back end, 16-level systems, FP32, one wavefront per instance.
// v_mfma_f32_16x16x4f32 operand layouts (lane L, k block kb):
//   A: A[L % 16][4 kb + L / 16]     B: B[4 kb + L / 16][L % 16]
//   C/D: lane L, item r holds D[4 (L / 16) + r][L % 16]
// Each lane owns the four density-matrix elements of its C/D slots.  H is
// rebuilt in A layout every stage from constant per-lane coefficients
// (H = Hd + e^{i phi} P1 + e^{-i phi} P2); rho enters as B through LDS, and
// X^dagger is read back transposed from LDS.  lds: 1024 floats per wave.
// ---------------------------------------------------------------------------
typedef float mm_f4 __attribute__((ext_vector_type(4)));

// LDS of the 16-level MFMA body: four 16-row arrays (stage re/im, X re/im)
// at a row pitch of 20 floats.  The lanes of a wavefront touch rows
// 4 rg + r (rg = lane / 16) at column lane % 16: at pitch 16 the four row
// groups fall on the same banks; at pitch 20 they are 16 banks apart, and
// the transposed 16-byte reads of X (row col, columns 4 rg .. 4 rg + 3) stay
// aligned and spread over every bank.
#define LB_MF_PITCH 20
#define LB_MF_LDS (4 * 16 * LB_MF_PITCH)

template <int NG>
__device__ __forceinline__ void lb_mfma16_body(const LbClass* __restrict__ C, long inst, bool active, int L,
                                               double kvi, double ph0, double dt, int n_steps, int avg_start,
                                               double* __restrict__ F_out, double* __restrict__ R_out,
                                               float* __restrict__ rho_out, float* __restrict__ lds) {
    static_assert(NG == 12, "the stacked MFMA product is laid out for 12 ground and 4 excited states");
    constexpr int N = 16, NE = 16 - NG;
    float* const sr = lds;
    float* const si = lds + 16 * LB_MF_PITCH;
    float* const xr = lds + 32 * LB_MF_PITCH;
    float* const xi = lds + 48 * LB_MF_PITCH;
    auto SP = [](int i, int j) { return i * LB_MF_PITCH + j; };
    const int col = L & 15, rg = L >> 4;

    // X = (C + C^dagger) rho with C = V in the (excited, ground) block,
    // V(t) = e^{i phi} A + e^{-i phi} B, in two parts on v_mfma_f32_16x16x4f32
    // (A operand: lane (rg, col) holds row col, k = rg; B and the result:
    // rows 4 kb + rg / 4 rg + r, column col):
    //  * excited rows, X_e = V rho_g (K = the 12 ground rows, k blocks 0..2):
    //    real and imaginary parts stacked in M, rows 0-3 Re, 4-7 Im:
    //      E = [Vr; Vi] rho_r + [-Vi; Vr] rho_i        (2 MFMA per k block)
    //  * ground rows, X_g = V^dagger rho_e (K = the 4 excited rows, block 3):
    //      Re = Vr^T rho_r + Vi^T rho_i,  Im = Vr^T rho_i - Vi^T rho_r  (4 MFMA)
    // 10 MFMA per product instead of the 16 of the dense 16 x 16 x 16 complex
    // product.  Each A entry is c P + s Q in (c, s) = (cos, sin) of the
    // phase; the lane's (P, Q) pairs:
    //   Vr = c (Ar + Br) + s (Bi - Ai),  Vi = c (Ai + Bi) + s (Ar - Br).
    float e1p[3], e1q[3], e2p[3], e2q[3];
#pragma unroll
    for (int kb = 0; kb < 3; kb++) {
        const int g = 4 * kb + rg;
        e1p[kb] = e1q[kb] = e2p[kb] = e2q[kb] = 0.f;
        if (col < 2 * NE) {
            const int e = col & 3;
            const float ar_ = (float)C->Ar[e][g], ai_ = (float)C->Ai[e][g];
            const float br_ = (float)C->Br[e][g], bi_ = (float)C->Bi[e][g];
            const float vrp = ar_ + br_, vrq = bi_ - ai_, vip = ai_ + bi_, viq = ar_ - br_;
            if (col < NE) {  // Re rows: Vr with rho_r, -Vi with rho_i
                e1p[kb] = vrp;
                e1q[kb] = vrq;
                e2p[kb] = -vip;
                e2q[kb] = -viq;
            } else {  // Im rows: Vi with rho_r, Vr with rho_i
                e1p[kb] = vip;
                e1q[kb] = viq;
                e2p[kb] = vrp;
                e2q[kb] = vrq;
            }
        }
    }
    float grp = 0.f, grq = 0.f, gip = 0.f, giq = 0.f;  // Vr[rg][col], Vi[rg][col]
    if (col < NG) {
        const float ar_ = (float)C->Ar[rg][col], ai_ = (float)C->Ai[rg][col];
        const float br_ = (float)C->Br[rg][col], bi_ = (float)C->Bi[rg][col];
        grp = ar_ + br_;
        grq = bi_ - ai_;
        gip = ai_ + bi_;
        giq = ar_ - br_;
    }
    struct AOps {
        float e1[3], e2[3], gr, gi;
    };
    auto build_a = [&](const float c, const float s, AOps& a) {
#pragma unroll
        for (int kb = 0; kb < 3; kb++) {
            a.e1[kb] = lb_fma(c, e1p[kb], s * e1q[kb]);
            a.e2[kb] = lb_fma(c, e2p[kb], s * e2q[kb]);
        }
        a.gr = lb_fma(c, grp, s * grq);
        a.gi = lb_fma(c, gip, s * giq);
    };

    // owned elements: (i, j) = (4 rg + r, col); rl, il: rounding error of the
    // state (compensated update), which also accumulates the step's RK4
    // increment h/6 (k1 + 2 k2 + 2 k3 + k4); dh: Hd_i - Hd_j formed in FP64, rounded once;
    // dec: (gam_i + gam_j) / 2.  Repopulation reaches only the diagonal ground
    // elements: at most one of a lane's four (r = rdi there), fed by
    // rep_d[e] rho_ee.
    float rr[4], ri[4], dec[4], rep_d[NE], rl[4], il[4], dh[4];
    const int rdi = (col < NG) ? col - 4 * rg : -1;
    const float gj = (col >= NG) ? (float)C->gam[col - NG] : 0.f;
#pragma unroll
    for (int e = 0; e < NE; e++) rep_d[e] = (col < NG && col / 4 == rg) ? (float)C->rep[col][e] : 0.f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int i = 4 * rg + r;
        rr[r] = (i == col && i < NG) ? 1.f / NG : 0.f;
        ri[r] = 0.f;
        rl[r] = il[r] = 0.f;
        dh[r] = (float)(C->Hd[i] - C->Hd[col]);
        dec[r] = 0.5f * (((i >= NG) ? (float)C->gam[i - NG] : 0.f) + gj);
    }

    // f(stage) -> (kr, ki) for the lane's elements; the stage is in sr/si
    // and, for the lane's own elements, in (qr, qi)
    auto rhs = [&](const AOps& a, const float (&qr)[4], const float (&qi)[4], float (&kr)[4], float (&ki)[4]) {
        mm_f4 E = {0.f, 0.f, 0.f, 0.f}, Gr = E, Gi = E;
#pragma unroll
        for (int kb = 0; kb < 3; kb++) {
            const float b_r = sr[SP(4 * kb + rg, col)];
            const float b_i = si[SP(4 * kb + rg, col)];
            E = __builtin_amdgcn_mfma_f32_16x16x4f32(a.e1[kb], b_r, E, 0, 0, 0);
            E = __builtin_amdgcn_mfma_f32_16x16x4f32(a.e2[kb], b_i, E, 0, 0, 0);
        }
        {
            const float b_r = sr[SP(NG + rg, col)];
            const float b_i = si[SP(NG + rg, col)];
            Gr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.gr, b_r, Gr, 0, 0, 0);
            Gi = __builtin_amdgcn_mfma_f32_16x16x4f32(a.gr, b_i, Gi, 0, 0, 0);
            Gr = __builtin_amdgcn_mfma_f32_16x16x4f32(a.gi, b_i, Gr, 0, 0, 0);
            Gi = __builtin_amdgcn_mfma_f32_16x16x4f32(-a.gi, b_r, Gi, 0, 0, 0);
        }
        // ground rows 4 rg + r from lanes rg < 3; excited rows NG + r: Re from
        // lanes rg = 0, Im from lanes rg = 1
        if (rg < 3) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                xr[SP(4 * rg + r, col)] = Gr[r];
                xi[SP(4 * rg + r, col)] = Gi[r];
            }
        }
        if (rg < 2) {
            float* const xe = rg == 0 ? xr : xi;
#pragma unroll
            for (int r = 0; r < 4; r++) xe[SP(NG + r, col)] = E[r];
        }
        __syncthreads();
        float rep = 0.f;
#pragma unroll
        for (int e = 0; e < NE; e++) rep = lb_fma(rep_d[e], sr[SP(NG + e, NG + e)], rep);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = 4 * rg + r;
            const int xdx = SP(i, col), xtd = SP(col, i);
            const float q_r = qr[r], q_i = qi[r];
            // -i (X_ij - conj(X_ji)) - i (Hd_i - Hd_j) rho_ij - dec rho_ij
            const float k_r = lb_fma(dh[r], q_i, lb_fma(-dec[r], q_r, xi[xdx] + xi[xtd]));
            kr[r] = (r == rdi) ? k_r + rep : k_r;
            ki[r] = lb_fma(-dh[r], q_r, lb_fma(-dec[r], q_i, xr[xtd] - xr[xdx]));
        }
        __syncthreads();
    };

    double fs = 0.0, rs = 0.0;
    const float h = (float)dt, h2 = (float)(0.5 * dt), h3 = (float)(dt / 3.0), h6 = (float)(dt / 6.0);
    // the phase advances by kvi dt per step: (cos, sin) at t + dt/2 and t + dt
    // by complex rotation of the value at t (FP64; the rotation's rounding
    // grows by ~1e-16 per step: ~3e-12 after 25,000 steps, far below the
    // FP32 state's resolution)
    double zc, zs, whc, whs, wc, ws;
    mm_cis(ph0, zc, zs);
    mm_cis(kvi * (0.5 * dt), whc, whs);
    mm_cis(kvi * dt, wc, ws);
    whc = lb_uniform(whc);
    whs = lb_uniform(whs);
    wc = lb_uniform(wc);
    ws = lb_uniform(ws);
    // one set of A operands is live at a time: the set at t serves stage 1,
    // then the set at t + dt/2 is built for stages 2 and 3, then the set at
    // t + dt for stage 4 and the next step's stage 1
    AOps a;
    build_a((float)zc, (float)zs, a);
    for (int step = 0; step < n_steps; step++) {
        zc = lb_uniform(zc);
        zs = lb_uniform(zs);
        float kr[4], ki[4], qr[4], qi[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            sr[SP(4 * rg + r, col)] = rr[r];
            si[SP(4 * rg + r, col)] = ri[r];
        }
        __syncthreads();
        rhs(a, rr, ri, kr, ki);
        build_a((float)(zc * whc - zs * whs), (float)(zs * whc + zc * whs), a);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            rl[r] = lb_fma(h6, kr[r], rl[r]);
            il[r] = lb_fma(h6, ki[r], il[r]);
            qr[r] = lb_fma(h2, kr[r], rr[r]);
            qi[r] = lb_fma(h2, ki[r], ri[r]);
            sr[SP(4 * rg + r, col)] = qr[r];
            si[SP(4 * rg + r, col)] = qi[r];
        }
        __syncthreads();
        rhs(a, qr, qi, kr, ki);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            rl[r] = lb_fma(h3, kr[r], rl[r]);
            il[r] = lb_fma(h3, ki[r], il[r]);
            qr[r] = lb_fma(h2, kr[r], rr[r]);
            qi[r] = lb_fma(h2, ki[r], ri[r]);
            sr[SP(4 * rg + r, col)] = qr[r];
            si[SP(4 * rg + r, col)] = qi[r];
        }
        __syncthreads();
        rhs(a, qr, qi, kr, ki);
#pragma unroll
        for (int r = 0; r < 4; r++) {
            rl[r] = lb_fma(h3, kr[r], rl[r]);
            il[r] = lb_fma(h3, ki[r], il[r]);
            qr[r] = lb_fma(h, kr[r], rr[r]);
            qi[r] = lb_fma(h, ki[r], ri[r]);
            sr[SP(4 * rg + r, col)] = qr[r];
            si[SP(4 * rg + r, col)] = qi[r];
        }
        __syncthreads();
        const double c1d = zc * wc - zs * ws, s1d = zs * wc + zc * ws;
        build_a((float)c1d, (float)s1d, a);
        rhs(a, qr, qi, kr, ki);
        // x + (lo + h/6 (k1 + 2 k2 + 2 k3 + k4)) with Fast2Sum: exact when
        // |x| >= the increment; otherwise the error it leaves in lo is below
        // ulp(x), far below the increment's own
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float incr = lb_fma(h6, kr[r], rl[r]), inci = lb_fma(h6, ki[r], il[r]);
            const float sr_ = rr[r] + incr, si_ = ri[r] + inci;
            rl[r] = incr - (sr_ - rr[r]);
            il[r] = inci - (si_ - ri[r]);
            rr[r] = sr_;
            ri[r] = si_;
        }
        if (step + 1 > avg_start) {
            // the weights are loaded here, not held across the step loop;
            // f = c U + s W (the weights' phase dependence taken out of the
            // sum).  Lanes rg < 3, col >= NG hold four (ground, excited)
            // coherences; lanes rg = 3 the excited row, its diagonal at r = col - NG.
            double U = 0.0, W = 0.0, rsum = 0.0;
            if (col >= NG) {
                if (rg < 3) {
                    const double* __restrict__ w = lb_fresh(C)->fw[col - NG][4 * rg];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const double x = (double)rr[r] + (double)rl[r], y = (double)ri[r] + (double)il[r];
                        U = __builtin_fma(x, w[4 * r + 3], __builtin_fma(y, w[4 * r], U));
                        W = __builtin_fma(x, w[4 * r + 2], __builtin_fma(-y, w[4 * r + 1], W));
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (r == col - NG) rsum = (double)rr[r] + (double)rl[r];
                }
            }
            // plain FP64 sums: over n <= 1e5 samples their rounding is below
            // n * 1.1e-16 of the summed magnitudes
            fs += __builtin_fma(c1d, U, s1d * W);
            rs += rsum;
        }
        zc = c1d;
        zs = s1d;
    }
    const double nsamp = (double)(n_steps - (avg_start < n_steps ? avg_start : n_steps));
    const double F = wave_sum(fs), R = wave_sum(rs);
    if (!active) return;
    if (L == 0) {
        F_out[inst] = nsamp > 0 ? C->force_unit * F / nsamp : 0.0;
        if (R_out) R_out[inst] = nsamp > 0 ? C->rate_unit * R / nsamp : 0.0;
    }
    if (rho_out) {
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int idx = (4 * rg + r) * N + col;
            rho_out[(inst * 256 + idx) * 2] = rr[r];
            rho_out[(inst * 256 + idx) * 2 + 1] = ri[r];
        }
    }
}

template <int NG>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_lindblad_rk4_mfma16(const LbClass* __restrict__ C, long n_inst,
                                                            const double* __restrict__ kv,
                                                            const double* __restrict__ phi0, double dt,
                                                            int n_steps, int avg_start, double* __restrict__ F_out,
                                                            double* __restrict__ R_out,
                                                            float* __restrict__ rho_out) {
    __shared__ __align__(16) float lds[LB_MF_LDS];
    const long inst = blockIdx.x;
    if (inst >= n_inst) return;
    lb_mfma16_body<NG>(C, inst, true, threadIdx.x, kv[inst], phi0[inst], dt, n_steps, avg_start, F_out, R_out,
                       rho_out, lds);
}

// ---------------------------------------------------------------------------
// Uneven batch in one launch: n2 instances of a large class (NG2 + NE2
// levels, BS2 lanes each, one block per instance) and n1 instances of a
// 16-level class (one wavefront each, BS2 / 64 per block; on MFMA for FP32,
// on the VALU body for FP64).  Large instances take the first blocks, so the
// long blocks start first.  Outputs as two separate launches would give.
// ---------------------------------------------------------------------------
struct LbUneven {
    const LbClass* c1;
    const LbClass* c2;
    long n1, n2;
    const double *kv1, *ph1, *kv2, *ph2;
    double *F1, *R1, *F2, *R2;
    void *rho1, *rho2;
};

template <typename T, int NG1, int NG2, int NE2, int BS2>
__global__ __launch_bounds__(BS2) __attribute__((amdgpu_waves_per_eu(4))) void k_lindblad_rk4_uneven(LbUneven U, double dt, int n_steps, int avg_start) {
    constexpr int W = BS2 / 64;
    constexpr size_t B1 = (sizeof(T) == 4) ? LB_MF_LDS * 4 : (size_t)lb_valu_lds<T, NG1, 16 - NG1>() * sizeof(T);
    constexpr size_t B2 = (size_t)lb_valu_lds<T, NG2, NE2>() * sizeof(T);
    constexpr size_t BYTES = (B1 * W > B2) ? B1 * W : B2;
    __shared__ __align__(16) unsigned char smem[BYTES];
    __shared__ double red[2 * W];
    const long b = blockIdx.x;
    if (b < U.n2) {
        lb_valu_body<T, NG2, NE2, BS2>(U.c2, b, true, threadIdx.x, U.kv2[b], U.ph2[b], dt, n_steps, avg_start, U.F2,
                                       U.R2, (T*)U.rho2, (T*)smem, red);
        return;
    }
    const int w = threadIdx.x >> 6, L = threadIdx.x & 63;
    const long want = (b - U.n2) * W + w;
    const bool active = want < U.n1;
    const long inst = active ? want : U.n1 - 1;  // a spare wave repeats the last instance, writes nothing
    if constexpr (sizeof(T) == 4)
        lb_mfma16_body<NG1>(U.c1, inst, active, L, U.kv1[inst], U.ph1[inst], dt, n_steps, avg_start, U.F1, U.R1,
                            (float*)U.rho1, (float*)smem + LB_MF_LDS * w);
    else
        lb_valu_body<T, NG1, 16 - NG1, 64>(U.c1, inst, active, L, U.kv1[inst], U.ph1[inst], dt, n_steps, avg_start,
                                           U.F1, U.R1, (T*)U.rho1, (T*)(smem + B1 * w), red + 0);
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int mm_lindblad_class_create(mm_lindblad_class** out, int n_ground, int n_excited, const double* Hd,
                                        const double* A_re, const double* A_im, const double* B_re,
                                        const double* B_im, const double* gam, const double* rep,
                                        double force_unit, double rate_unit) {
    mm_clear_error();
    if (!out) return mm_fail(MM_EINVAL, "lindblad_class: null output");
    *out = nullptr;
    if (n_ground < 1 || n_ground > LB_MAXG || n_excited < 1 || n_excited > LB_MAXE)
        return mm_fail(MM_EINVAL, "lindblad_class: n_ground in [1,%d], n_excited in [1,%d]", LB_MAXG, LB_MAXE);
    if (!Hd || !A_re || !A_im || !B_re || !B_im || !gam || !rep) return mm_fail(MM_EINVAL, "lindblad_class: null table");
    mm_lindblad_class* c = new mm_lindblad_class;
    c->lanes[0] = c->lanes[1] = 0;
    memset(&c->host, 0, sizeof(LbClass));
    LbClass& L = c->host;
    L.ng = n_ground;
    L.ne = n_excited;
    for (int i = 0; i < n_ground + n_excited; i++) L.Hd[i] = Hd[i];
    for (int e = 0; e < n_excited; e++) {
        L.gam[e] = gam[e];
        for (int g = 0; g < n_ground; g++) {
            L.Ar[e][g] = A_re[e * n_ground + g];
            L.Ai[e][g] = A_im[e * n_ground + g];
            L.Br[e][g] = B_re[e * n_ground + g];
            L.Bi[e][g] = B_im[e * n_ground + g];
        }
    }
    for (int g = 0; g < n_ground; g++)
        for (int e = 0; e < n_excited; e++) L.rep[g][e] = rep[g * n_excited + e];
    L.force_unit = force_unit;
    L.rate_unit = rate_unit;
    for (int e = 0; e < n_excited; e++)
        for (int g = 0; g < n_ground; g++) {
            L.fw[e][g][0] = L.Ar[e][g] - L.Br[e][g];
            L.fw[e][g][1] = L.Ai[e][g] + L.Bi[e][g];
            L.fw[e][g][2] = L.Ar[e][g] + L.Br[e][g];
            L.fw[e][g][3] = L.Ai[e][g] - L.Bi[e][g];
        }
    if (hipMalloc((void**)&c->dev, sizeof(LbClass)) != hipSuccess) {
        delete c;
        return mm_hip_fail("hipMalloc(lindblad class)");
    }
    if (hipMemcpy(c->dev, &c->host, sizeof(LbClass), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(c->dev);
        delete c;
        return mm_hip_fail("hipMemcpy(lindblad class)");
    }
    *out = c;
    return MM_OK;
}

extern "C" int mm_lindblad_class_destroy(mm_lindblad_class* c) {
    mm_clear_error();
    if (!c) return MM_OK;
    int rc = MM_OK;
    if (c->dev && hipFree(c->dev) != hipSuccess) rc = mm_hip_fail("hipFree(lindblad class)");
    delete c;
    return rc;
}

// Lanes per instance of the VALU back end: the class's setting, else one
// wavefront up to 16 levels and four above (fewer elements per lane).
static int lb_lanes(const mm_lindblad_class* c, int precision) {
    const int set = c->lanes[precision == 0 ? 0 : 1];
    if (set) return set;
    const int n = c->host.ng + c->host.ne;
    return n * n <= 256 ? 64 : 256;
}

template <typename T, int NG, int NE>
static int launch_valu(const mm_lindblad_class* c, int lanes, long n, const double* kv, const double* phi0,
                       double dt, int n_steps, int avg_start, double* F, double* R, void* rho, hipStream_t s) {
    if (lanes == 64)
        hipLaunchKernelGGL((k_lindblad_rk4_valu<T, NG, NE, 64>), dim3((unsigned)n), dim3(64), 0, s, c->dev, n, kv,
                           phi0, dt, n_steps, avg_start, F, R, (T*)rho);
    else if (lanes == 128)
        hipLaunchKernelGGL((k_lindblad_rk4_valu<T, NG, NE, 128>), dim3((unsigned)n), dim3(128), 0, s, c->dev, n, kv,
                           phi0, dt, n_steps, avg_start, F, R, (T*)rho);
    else
        hipLaunchKernelGGL((k_lindblad_rk4_valu<T, NG, NE, 256>), dim3((unsigned)n), dim3(256), 0, s, c->dev, n, kv,
                           phi0, dt, n_steps, avg_start, F, R, (T*)rho);
    return mm_check_launch("k_lindblad_rk4_valu");
}

extern "C" int mm_lindblad_class_set_lanes(mm_lindblad_class* c, int precision, int lanes) {
    mm_clear_error();
    if (!c) return mm_fail(MM_EINVAL, "lindblad_set_lanes: null class");
    if (precision != 0 && precision != 1) return mm_fail(MM_EINVAL, "lindblad_set_lanes: precision 0 or 1");
    if (lanes != 0 && lanes != 64 && lanes != 128 && lanes != 256)
        return mm_fail(MM_EINVAL, "lindblad_set_lanes: lanes is 0 (default), 64, 128 or 256");
    c->lanes[precision] = lanes;
    return MM_OK;
}

extern "C" int mm_lindblad_class_get_lanes(const mm_lindblad_class* c, int precision) {
    mm_clear_error();
    if (!c || (precision != 0 && precision != 1)) return -mm_fail(MM_EINVAL, "lindblad_get_lanes: bad arguments");
    return lb_lanes(c, precision);
}

// ---------------------------------------------------------------------------
// Augmented Liouvillian for the steady-state solve (natural units):
// L = -i (I (x) H - H^T (x) I) + dissipator, column-major vec (index i + N j
// for rho_ij), as build_liouvillian_fast (molmot/obe/obe_subdoppler.py:107-141),
// with its last row replaced by the trace constraint as steady_state_batch
// does (molmot_cuda/lindblad_cuda.py:224-232).  One thread per element;
// output complex128 interleaved, (n_inst, N^2, N^2) row-major.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void lb_H(const LbClass* __restrict__ C, int N, int i, int k, double c, double s,
                                     double& hr, double& hi) {
    const int NG = C->ng;
    hr = (i == k) ? C->Hd[i] : 0.0;
    hi = 0.0;
    if (i >= NG && k < NG) {  // (e, g): e^{i phi} A + e^{-i phi} B
        const int e = i - NG, g = k;
        hr += c * C->Ar[e][g] - s * C->Ai[e][g] + c * C->Br[e][g] + s * C->Bi[e][g];
        hi += s * C->Ar[e][g] + c * C->Ai[e][g] - s * C->Br[e][g] + c * C->Bi[e][g];
    } else if (i < NG && k >= NG) {  // conjugate of the (k, i) element
        const int e = k - NG, g = i;
        hr += c * C->Ar[e][g] - s * C->Ai[e][g] + c * C->Br[e][g] + s * C->Bi[e][g];
        hi -= s * C->Ar[e][g] + c * C->Ai[e][g] - s * C->Br[e][g] + c * C->Bi[e][g];
    }
}

__global__ void k_lindblad_liouvillian(const LbClass* __restrict__ C, long n_inst, const double* __restrict__ phi0,
                                       int trace_row, double* __restrict__ out) {
    const int N = C->ng + C->ne;
    const long N2 = (long)N * N;
    const long el = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long inst = blockIdx.y;
    if (el >= N2 * N2 || inst >= n_inst) return;
    const long r = el / N2, col = el - (el / N2) * N2;
    const int a = (int)(r % N), b = (int)(r / N), cc = (int)(col % N), d = (int)(col / N);
    double vr = 0.0, vi = 0.0;
    if (trace_row && r == N2 - 1) {
        vr = (cc == d) ? 1.0 : 0.0;
    } else {
        const double ph = phi0[inst];
        const double cph = cos(ph), sph = sin(ph);
        double hr, hi;
        // -i (delta_bd H[a][cc] - delta_{a,cc} H[d][b])
        if (b == d) {
            lb_H(C, N, a, cc, cph, sph, hr, hi);
            vr += hi;
            vi -= hr;
        }
        if (a == cc) {
            lb_H(C, N, d, b, cph, sph, hr, hi);
            vr -= hi;
            vi += hr;
        }
        const int NG = C->ng;
        if (r == col) vr -= 0.5 * ((a >= NG ? C->gam[a - NG] : 0.0) + (b >= NG ? C->gam[b - NG] : 0.0));
        if (a == b && cc == d && a < NG && cc >= NG) vr += C->rep[a][cc - NG];
    }
    out[2 * (inst * N2 * N2 + el)] = vr;
    out[2 * (inst * N2 * N2 + el) + 1] = vi;
}

extern "C" int mm_lindblad_liouvillian(const mm_lindblad_class* c, long n_inst, const double* phi0, int trace_row,
                                       double* out, void* stream) {
    mm_clear_error();
    if (!c) return mm_fail(MM_EINVAL, "lindblad_liouvillian: null class");
    if (n_inst < 0) return mm_fail(MM_EINVAL, "lindblad_liouvillian: negative size");
    if (n_inst == 0) return MM_OK;
    if (!phi0 || !out) return mm_fail(MM_EINVAL, "lindblad_liouvillian: null pointer");
    if (n_inst > 65535) return mm_fail(MM_EINVAL, "lindblad_liouvillian: at most 65535 instances per call");
    const long N = c->host.ng + c->host.ne;
    const long n_el = N * N * N * N;
    hipLaunchKernelGGL(k_lindblad_liouvillian, dim3((unsigned)((n_el + 255) / 256), (unsigned)n_inst), dim3(256), 0,
                       (hipStream_t)stream, c->dev, n_inst, phi0, trace_row, out);
    return mm_check_launch("k_lindblad_liouvillian");
}

// (n_ground, n_excited) classes with compiled propagators
#define LB_SIZE_LIST(X) \
    X(12, 4)            \
    X(36, 4)            \
    X(1, 1)             \
    X(2, 2)

extern "C" int mm_lindblad_rk4(const mm_lindblad_class* c, long n_inst, const double* kv, const double* phi0,
                               double dt, int n_steps, int avg_start, int precision, double* F, double* R,
                               void* rho_final, void* stream) {
    mm_clear_error();
    if (!c) return mm_fail(MM_EINVAL, "lindblad_rk4: null class");
    if (n_inst < 0 || n_steps < 0) return mm_fail(MM_EINVAL, "lindblad_rk4: negative size");
    if (n_inst == 0) return MM_OK;
    if (!kv || !phi0 || !F) return mm_fail(MM_EINVAL, "lindblad_rk4: null pointer");
    if (n_inst > 0x7fffffffL) return mm_fail(MM_EINVAL, "lindblad_rk4: too many instances for one launch");
    const int ng = c->host.ng, ne = c->host.ne;
    hipStream_t s = (hipStream_t)stream;
    if (precision == 2) {
        if (ng + ne != 16) return mm_fail(MM_EUNSUPPORTED, "lindblad_rk4: the MFMA path takes 16-level systems");
        if (ng == 12) {
            hipLaunchKernelGGL((k_lindblad_rk4_mfma16<12>), dim3((unsigned)n_inst), dim3(64), 0, s, c->dev, n_inst,
                               kv, phi0, dt, n_steps, avg_start, F, R, (float*)rho_final);
            return mm_check_launch("k_lindblad_rk4_mfma16");
        }
        return mm_fail(MM_EUNSUPPORTED, "lindblad_rk4: MFMA path compiled for 12+4 only");
    }
    if (precision != 0 && precision != 1) return mm_fail(MM_EINVAL, "lindblad_rk4: precision must be 0, 1 or 2");
    const int lanes = lb_lanes(c, precision);
#define LB_CASE(G, E)                                                                                        \
    if (ng == G && ne == E)                                                                                  \
        return precision == 0 ? launch_valu<double, G, E>(c, lanes, n_inst, kv, phi0, dt, n_steps, avg_start, F, \
                                                          R, rho_final, s)                                   \
                              : launch_valu<float, G, E>(c, lanes, n_inst, kv, phi0, dt, n_steps, avg_start, F,  \
                                                         R, rho_final, s);
    LB_SIZE_LIST(LB_CASE)
#undef LB_CASE
    return mm_fail(MM_EUNSUPPORTED, "lindblad_rk4: no propagator for (%d, %d)", ng, ne);
}

extern "C" int mm_lindblad_rk4_uneven(const mm_lindblad_class* c16, long n16, const double* kv16,
                                      const double* phi16, double* F16, double* R16, void* rho16,
                                      const mm_lindblad_class* cbig, long nbig, const double* kvbig,
                                      const double* phibig, double* Fbig, double* Rbig, void* rhobig, double dt,
                                      int n_steps, int avg_start, int precision, void* stream) {
    mm_clear_error();
    if (!c16 || !cbig) return mm_fail(MM_EINVAL, "lindblad_rk4_uneven: null class");
    if (n16 < 0 || nbig < 0 || n_steps < 0) return mm_fail(MM_EINVAL, "lindblad_rk4_uneven: negative size");
    if ((n16 && (!kv16 || !phi16 || !F16)) || (nbig && (!kvbig || !phibig || !Fbig)))
        return mm_fail(MM_EINVAL, "lindblad_rk4_uneven: null pointer");
    if (c16->host.ng != 12 || c16->host.ne != 4 || cbig->host.ng != 36 || cbig->host.ne != 4)
        return mm_fail(MM_EUNSUPPORTED, "lindblad_rk4_uneven: compiled for a 12+4 and a 36+4 class");
    if (precision != 0 && precision != 2)
        return mm_fail(MM_EINVAL, "lindblad_rk4_uneven: precision 0 (FP64) or 2 (FP32, MFMA for the 16-level part)");
    const long blocks = nbig + (n16 + 3) / 4;
    if (blocks == 0) return MM_OK;
    if (blocks > 0x7fffffffL) return mm_fail(MM_EINVAL, "lindblad_rk4_uneven: too many instances for one launch");
    LbUneven U;
    U.c1 = c16->dev;
    U.c2 = cbig->dev;
    U.n1 = n16;
    U.n2 = nbig;
    U.kv1 = kv16;
    U.ph1 = phi16;
    U.kv2 = kvbig;
    U.ph2 = phibig;
    U.F1 = F16;
    U.R1 = R16;
    U.F2 = Fbig;
    U.R2 = Rbig;
    U.rho1 = rho16;
    U.rho2 = rhobig;
    hipStream_t s = (hipStream_t)stream;
    if (precision == 0)
        hipLaunchKernelGGL((k_lindblad_rk4_uneven<double, 12, 36, 4, 256>), dim3((unsigned)blocks), dim3(256), 0, s, U,
                           dt, n_steps, avg_start);
    else
        hipLaunchKernelGGL((k_lindblad_rk4_uneven<float, 12, 36, 4, 256>), dim3((unsigned)blocks), dim3(256), 0, s, U,
                           dt, n_steps, avg_start);
    return mm_check_launch("k_lindblad_rk4_uneven");
}
