// Monte Carlo wavefunction (quantum jump) ensemble on gfx950.
//
// Physics follows SSESolver (molmot/obe/stochastic.py:579-1067) and the
// device reference (molmot_cuda/stochastic_cuda.py:404-756): RK4 on the
// unnormalised amplitudes c under the six-beam field with Gaussian profiles,
// non-Hermitian decay -c_e/2 and off-diagonal ground Zeeman coupling; a jump
// when 1 - |c|^2 exceeds -ln U, channel (e, g, q) chosen with probability
// br * |c_e|^2 by a cumulative search, then c = |g>, isotropic recoil, optional
// diffusion kick and renormalisation; (r, v) integrated with the same RK4
// using the dipole force.
//
// Layout and precision:
//  * one lane per trajectory, 64 trajectories per wavefront, the whole step
//    loop inside one launch; every table is read with wave-uniform indices
//    (scalar loads);
//  * amplitudes and the right-hand side in T (float or double); r, v, t, the
//    norm, the jump threshold, the channel search and the accumulated force in
//    double;
//  * the laser sum S_{b,q}(t) = sum_f G_f conj(eps_{b,f,q}) exp(i w_f t) does
//    not depend on the trajectory: k_mcwf_fields tabulates it for every stage
//    time (t_k and t_k + dt/2) once, and the step kernel multiplies it by the
//    per-lane Gaussian amplitude and exp(-+i r);
//  * energies and laser frequencies are measured from manifold references
//    (ground mean, excited mean; lasers from their difference), so every FP64
//    phase argument stays below ~1e4 rad; the coupling phases
//    w_f - w_e + w_g are unchanged;
//  * the Zeeman term is applied directly to c: its interaction-picture phases
//    exp(-i w_i t) exp(+i w_j t) cancel against psi_int and the back-transform.
//
// Random numbers: Philox4x32-10, key (seed_lo, seed_hi), counter
// (block, stream, trajectory_lo, trajectory_hi); block 0 = first threshold,
// jump m (1-based): block 3m-2 (channel, next threshold), 3m-1 (recoil
// direction), 3m (diffusion signs); renormalisation redraw r uses block r of
// stream 1.  u = ((a >> 5) 2^26 + (b >> 6) + 1/2) 2^-53 in (0, 1).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "mm_errors.h"
#include "mm_trig.h"
#include "philox.h"

#define MC_MAXG 12
#define MC_MAXE 4
#define MC_MAXN (MC_MAXG + MC_MAXE)
#define MC_MAXF 8

template <typename T>
struct McTab {
    T d[MC_MAXG][MC_MAXE][3];              // |d_{g,e,q}|
    T zr[MC_MAXG][MC_MAXG][3];             // ground Zeeman blocks (x, y, z), real part
    T zi[MC_MAXG][MC_MAXG][3];             // imaginary part
};

struct McProblem {
    int ng, ne, nf, add_kick;
    double wbase[MC_MAXN];                 // shifted energies (natural units)
    double zgd[MC_MAXG][3];                // Re diag of the ground Zeeman blocks
    double zde[MC_MAXE];                   // excited Bz diagonal
    double br[MC_MAXE][MC_MAXG][3];        // branching ratios
    double amp_re[6][MC_MAXF][3], amp_im[6][MC_MAXF][3];  // G_f conj(eps_{b,f,q})
    double wl[MC_MAXF];                    // shifted laser frequencies
    double B_gradient, r_nat, denom, mass_nat, v_recoil;
    double diffusion[3];
    McTab<float> tf;
    McTab<double> td;
};

struct mm_mcwf_problem {
    McProblem host;
    McProblem* dev;
};

template <typename T>
struct C2 {
    T re, im;
};

template <typename T>
__device__ __forceinline__ const McTab<T>& tab_of(const McProblem* P);
template <>
__device__ __forceinline__ const McTab<float>& tab_of<float>(const McProblem* P) {
    return P->tf;
}
template <>
__device__ __forceinline__ const McTab<double>& tab_of<double>(const McProblem* P) {
    return P->td;
}

// Gaussian amplitude profile sqrt(exp(-r_perp^2 / denom)) (stochastic.py:233-241)
template <typename T>
__device__ __forceinline__ T mc_gauss(double rp, double denom);
template <>
__device__ __forceinline__ double mc_gauss<double>(double rp, double denom) {
    return sqrt(exp(-rp / denom));
}
template <>
__device__ __forceinline__ float mc_gauss<float>(double rp, double denom) {
    return __expf(-0.5f * (float)(rp / denom));
}

// Table pointers in the constant address space: loads through them are
// scalar loads.  The empty asm statements below hide the pointer from
// loop-invariant code motion, so that one block of table entries is loaded
// where it is used instead of all blocks being hoisted (and spilled) at the
// top of the step loop; the "after" form also orders the block after the
// computation of x, which keeps unrolled iterations from being interleaved
// into one huge live set.
#ifdef __HIP_DEVICE_COMPILE__
#define MC_CONST __attribute__((address_space(4)))
#else
#define MC_CONST
#endif
typedef const MC_CONST McProblem* McCPtr;

__device__ __forceinline__ McCPtr mc_const(const McProblem* p) { return (McCPtr)(unsigned long long)p; }

__device__ __forceinline__ McCPtr mc_fresh(McCPtr p) {
    asm volatile("" : "+s"(p));
    return p;
}

template <typename T>
__device__ __forceinline__ McCPtr mc_fresh_after(McCPtr p, T x) {
    asm volatile("" : "+s"(p) : "v"(x));
    return p;
}

template <typename T>
__device__ __forceinline__ const MC_CONST McTab<T>& tab_c(McCPtr P);
template <>
__device__ __forceinline__ const MC_CONST McTab<float>& tab_c<float>(McCPtr P) {
    return P->tf;
}
template <>
__device__ __forceinline__ const MC_CONST McTab<double>& tab_c<double>(McCPtr P) {
    return P->td;
}

__device__ __forceinline__ double mc_uniform(uint32_t a, uint32_t b) {
    return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6) + 0.5) * 1.1102230246251565e-16;
}

// Jump threshold from a uniform u: the reference compares 1 - |psi|^2 with
// -ln u (stochastic.py:742-746, stochastic_cuda.py:600), under which a
// threshold above 1 never fires; rule 1 uses 1 - u, i.e. a jump when
// |psi|^2 < u, the exact waiting-time distribution of the unravelling.
__device__ __forceinline__ double mc_threshold(double u, int rule) { return rule ? 1.0 - u : -log(u); }

__device__ __forceinline__ mm_u32x4 mc_draw(uint64_t seed, long traj, uint32_t block, uint32_t stream) {
    mm_u32x4 c;
    c.x = block;
    c.y = stream;
    c.z = (uint32_t)((uint64_t)traj & 0xffffffffu);
    c.w = (uint32_t)((uint64_t)traj >> 32);
    return mm_philox4x32_10(c, (uint32_t)(seed & 0xffffffffu), (uint32_t)(seed >> 32));
}

// a*b + c with one rounding.  The library is compiled with contraction off
// (the rate-equation detunings need the reference's two roundings), so the
// amplitude arithmetic writes its fused products out: v_fma_f32 / v_fma_f64
// instead of a multiply and an add.
__device__ __forceinline__ float mc_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double mc_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }

// a b
template <typename T>
__device__ __forceinline__ C2<T> mc_mul(C2<T> a, C2<T> b) {
    return {mc_fma(a.re, b.re, -(a.im * b.im)), mc_fma(a.re, b.im, a.im * b.re)};
}
// acc + a b
template <typename T>
__device__ __forceinline__ C2<T> mc_mac(C2<T> acc, C2<T> a, C2<T> b) {
    return {mc_fma(a.re, b.re, mc_fma(-a.im, b.im, acc.re)), mc_fma(a.re, b.im, mc_fma(a.im, b.re, acc.im))};
}
// acc + conj(a) b
template <typename T>
__device__ __forceinline__ C2<T> mc_cmac(C2<T> acc, C2<T> a, C2<T> b) {
    return {mc_fma(a.re, b.re, mc_fma(a.im, b.im, acc.re)), mc_fma(a.re, b.im, mc_fma(-a.im, b.re, acc.im))};
}
// -i x conj(p): back from the interaction picture of the state whose phase
// factor is p = exp(-i w t)
template <typename T>
__device__ __forceinline__ C2<T> mc_back(C2<T> x, C2<T> p) {
    return {mc_fma(x.im, p.re, -(x.re * p.im)), mc_fma(-x.re, p.re, -(x.im * p.im))};
}

// Right-hand side at one stage: dc/dt for unnormalised c and the force
// (natural units, unnormalised).  S points at the 18 complex laser sums of
// this stage time.
template <typename T, int NG, int NE>
__device__ __forceinline__ void mc_rhs(McCPtr P, const C2<T>* __restrict__ S, double t,
                                       const C2<T> (&c)[NG + NE], const double (&r)[3], C2<T> (&dc)[NG + NE],
                                       double (&F)[3]) {
    const McCPtr PC = mc_fresh(P);
    const MC_CONST C2<T>* SC = (const MC_CONST C2<T>*)(unsigned long long)S;
    // Quadrupole field in Gauss (stochastic.py:526-549)
    const double Bx = -PC->B_gradient * (r[0] * PC->r_nat) * 0.5 * 1e4;
    const double By = -PC->B_gradient * (r[1] * PC->r_nat) * 0.5 * 1e4;
    const double Bz = PC->B_gradient * (r[2] * PC->r_nat) * 1e4;
    // interaction-picture phases exp(-i w_n t) with the diagonal Zeeman shift
    // exp(-i w_n t) for the excited states here; each ground state's phase is
    // formed inside its own coupling iteration below
    C2<T> phe[NE], pe[NE];
#pragma unroll
    for (int e = 0; e < NE; e++) {
        const double w = mc_fma(Bz, PC->zde[e], PC->wbase[NG + e]);
        T cs, sn;
        mm_cis(w * t, cs, sn);
        phe[e].re = cs;
        phe[e].im = -sn;
        pe[e] = mc_mul(c[NG + e], phe[e]);
    }
    // fields: E_{b,q} = gauss_axis exp(-+i r_axis) S_{b,q}
    C2<T> Et[3], Ed[3][3];
#pragma unroll
    for (int q = 0; q < 3; q++) Et[q].re = Et[q].im = T(0);
    const double rp[3] = {r[1] * r[1] + r[2] * r[2], r[0] * r[0] + r[2] * r[2], r[0] * r[0] + r[1] * r[1]};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const T ga = mc_gauss<T>(rp[k], PC->denom);
        T cs, sn;
        mm_cis(r[k], cs, sn);
        const C2<T> af = {ga * cs, -ga * sn}, ab = {ga * cs, ga * sn};
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const C2<T> ef = mc_mul(af, SC[k * 3 + q]), eb = mc_mul(ab, SC[(k + 3) * 3 + q]);
            Et[q].re += ef.re + eb.re;
            Et[q].im += ef.im + eb.im;
            Ed[k][q].re = ef.re - eb.re;
            Ed[k][q].im = ef.im - eb.im;
        }
    }
    // dipole contractions in one pass over ground states
    C2<T> pqe[NE][3], dexp[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        dexp[q].re = dexp[q].im = T(0);
#pragma unroll
        for (int e = 0; e < NE; e++) pqe[e][q].re = pqe[e][q].im = T(0);
    }
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const McCPtr Pg = mc_fresh_after(PC, g ? dc[g - 1].re : pe[0].re);
        const MC_CONST McTab<T>& Tb = tab_c<T>(Pg);
        C2<T> phg;
        {
            const double w =
                mc_fma(Bz, Pg->zgd[g][2], mc_fma(By, Pg->zgd[g][1], mc_fma(Bx, Pg->zgd[g][0], Pg->wbase[g])));
            T cs, sn;
            mm_cis(w * t, cs, sn);
            phg.re = cs;
            phg.im = -sn;
        }
        const C2<T> pg = mc_mul(c[g], phg);
        C2<T> pqg[3];
        T vr = T(0), vi = T(0);
#pragma unroll
        for (int q = 0; q < 3; q++) {
            pqg[q].re = pqg[q].im = T(0);
#pragma unroll
            for (int e = 0; e < NE; e++) {
                const T dd = Tb.d[g][e][q];
                pqg[q].re = mc_fma(dd, pe[e].re, pqg[q].re);
                pqg[q].im = mc_fma(dd, pe[e].im, pqg[q].im);
                pqe[e][q].re = mc_fma(dd, pg.re, pqe[e][q].re);
                pqe[e][q].im = mc_fma(dd, pg.im, pqe[e][q].im);
            }
            // d_exp[q] += conj(pint_g) pqg[q]
            dexp[q] = mc_cmac(dexp[q], pg, pqg[q]);
            // sum_q Et[q] pqg[q]
            const C2<T> v = mc_mac(C2<T>{vr, vi}, Et[q], pqg[q]);
            vr = v.re;
            vi = v.im;
        }
        // dpsi_int_g = -i sum_q Et[q] pqg[q], back to c with conj(ph_g)
        dc[g] = mc_back(C2<T>{vr, vi}, phg);
    }
#pragma unroll
    for (int e = 0; e < NE; e++) {
        // sum_q conj(Et[q]) pqe[e][q] - (i/2) p_e
        C2<T> v = {T(0.5) * pe[e].im, -T(0.5) * pe[e].re};
#pragma unroll
        for (int q = 0; q < 3; q++) v = mc_cmac(v, Et[q], pqe[e][q]);
        dc[NG + e] = mc_back(v, phe[e]);
    }
    // off-diagonal ground Zeeman coupling on c: dc_g += -i sum_{j != g} H_gj c_j
    const T bx = (T)Bx, by = (T)By, bz = (T)Bz;
#pragma unroll
    for (int g = 0; g < NG; g++) {
        const MC_CONST McTab<T>& Tb = tab_c<T>(mc_fresh_after(PC, dc[g].re));
#pragma unroll
        for (int j = g + 1; j < NG; j++) {
            const T hr = mc_fma(bx, Tb.zr[g][j][0], mc_fma(by, Tb.zr[g][j][1], bz * Tb.zr[g][j][2]));
            const T hi = mc_fma(bx, Tb.zi[g][j][0], mc_fma(by, Tb.zi[g][j][1], bz * Tb.zi[g][j][2]));
            // -i H_gj c_j
            dc[g].re = mc_fma(hr, c[j].im, mc_fma(hi, c[j].re, dc[g].re));
            dc[g].im = mc_fma(-hr, c[j].re, mc_fma(hi, c[j].im, dc[g].im));
            // -i conj(H_gj) c_g into row j
            dc[j].re = mc_fma(hr, c[g].im, mc_fma(-hi, c[g].re, dc[j].re));
            dc[j].im = mc_fma(-hr, c[g].re, mc_fma(-hi, c[g].im, dc[j].im));
        }
    }
    // F_k = -2 sum_q Im(d_exp[q] Ed[k][q])
#pragma unroll
    for (int k = 0; k < 3; k++) {
        T s = T(0);
#pragma unroll
        for (int q = 0; q < 3; q++) s = mc_fma(dexp[q].re, Ed[k][q].im, mc_fma(dexp[q].im, Ed[k][q].re, s));
        F[k] = -2.0 * (double)s;
    }
}

struct McRun {
    long n;
    double dt, r_escape2;
    int n_steps, save_every, n_save, renorm_interval, avg_start, jump_cap, jump_rule;
    uint64_t seed;
    const double *r0, *v0, *psi0;
    const void* S;
    double *r_out, *v_out, *pop_out, *f_out, *fsum, *psi_final, *r_final, *v_final, *jump_t;
    int *n_scatters, *escaped, *n_saved, *n_avg, *jump_ch;
};

template <typename T, int NG, int NE>
__global__ __launch_bounds__(64) void k_mcwf(const McProblem* __restrict__ Pg, McRun R) {
    const McCPtr P = mc_const(Pg);
    constexpr int N = NG + NE;
    // RK4 accumulator in a thread-private LDS column (stride 64: no bank
    // conflicts), which keeps the register footprint of the stage loop down
    __shared__ T acc_lds[2 * (NG + NE)][64];
    T* const acc = &acc_lds[0][threadIdx.x];
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= R.n) return;
    const C2<T>* __restrict__ S = (const C2<T>*)R.S;
    C2<T> y[N];
    double r[3], v[3];
    if (R.psi0) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            y[i].re = (T)R.psi0[(tid * N + i) * 2];
            y[i].im = (T)R.psi0[(tid * N + i) * 2 + 1];
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; i++) y[i].re = y[i].im = T(0);
        y[0].re = T(1);
    }
#pragma unroll
    for (int i = 0; i < 3; i++) {
        r[i] = R.r0[tid * 3 + i];
        v[i] = R.v0[tid * 3 + i];
    }
    mm_u32x4 w0 = mc_draw(R.seed, tid, 0u, 0u);
    double thr = mc_threshold(mc_uniform(w0.x, w0.y), R.jump_rule);
    double t = 0.0, last = 0.0;
    double fs[3] = {0.0, 0.0, 0.0};
    int n_sc = 0, esc = 0, si = 0, n_av = 0, n_logged = 0;
    uint32_t n_jump = 0, n_ren = 0;
    const double dt = R.dt, hdt = 0.5 * R.dt, dt6 = R.dt / 6.0;
    const double m = P->mass_nat;
    for (int step = 0;; ++step) {
        const bool last_step = step == R.n_steps;
        const bool escape_now =
            !last_step && R.r_escape2 > 0.0 && (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) > R.r_escape2;
        double ns0 = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) ns0 += (double)y[i].re * (double)y[i].re + (double)y[i].im * (double)y[i].im;
        // the four RK4 stages share one code instance; stage 1's force also
        // feeds the save point and the force average
        C2<T> k[N], tmp[N];
        double F[3], rs[3], vs[3], ar[3], av[3];
#pragma unroll
        for (int i = 0; i < N; i++) tmp[i] = y[i];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            rs[i] = r[i];
            vs[i] = v[i];
        }
        bool stop = false;
#pragma unroll 1
        for (int st = 0; st < 4; ++st) {
            const long tau = 2L * step + (st == 0 ? 0 : (st == 3 ? 2 : 1));
            mc_rhs<T, NG, NE>(P, S + tau * 18, st == 0 ? t : (st == 3 ? t + dt : t + hdt), tmp, rs, k, F);
            if (st == 0) {
                if (step % R.save_every == 0 && si < R.n_save) {
                    const double inv = ns0 > 1e-30 ? 1.0 / ns0 : 1.0;
                    const long o = tid * R.n_save + si;
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        if (R.r_out) R.r_out[o * 3 + i] = r[i];
                        if (R.v_out) R.v_out[o * 3 + i] = v[i];
                        if (R.f_out) R.f_out[o * 3 + i] = F[i] * inv;
                    }
                    if (R.pop_out) {
#pragma unroll
                        for (int i = 0; i < N; i++)
                            R.pop_out[o * N + i] =
                                ((double)y[i].re * (double)y[i].re + (double)y[i].im * (double)y[i].im) * inv;
                    }
                    si++;
                }
                if (last_step || escape_now) {
                    esc = escape_now ? 1 : 0;
                    stop = true;
                    break;
                }
                if (step >= R.avg_start) {
#pragma unroll
                    for (int i = 0; i < 3; i++) fs[i] += F[i] / ns0;
                    n_av++;
                }
            }
            if (st < 3) {
                const double hn = st == 2 ? dt : hdt;
#pragma unroll
                for (int i = 0; i < N; i++) {
                    const T are = st == 0 ? k[i].re : mc_fma(T(2), k[i].re, acc[(2 * i) * 64]);
                    const T aim = st == 0 ? k[i].im : mc_fma(T(2), k[i].im, acc[(2 * i + 1) * 64]);
                    acc[(2 * i) * 64] = are;
                    acc[(2 * i + 1) * 64] = aim;
                    tmp[i].re = mc_fma((T)hn, k[i].re, y[i].re);
                    tmp[i].im = mc_fma((T)hn, k[i].im, y[i].im);
                }
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    const double kr = vs[i], kv = F[i] / m;
                    ar[i] = st == 0 ? kr : ar[i] + 2.0 * kr;
                    av[i] = st == 0 ? kv : av[i] + 2.0 * kv;
                    rs[i] = r[i] + hn * kr;
                    vs[i] = v[i] + hn * kv;
                }
            } else {
#pragma unroll
                for (int i = 0; i < N; i++) {
                    y[i].re = mc_fma((T)dt6, acc[(2 * i) * 64] + k[i].re, y[i].re);
                    y[i].im = mc_fma((T)dt6, acc[(2 * i + 1) * 64] + k[i].im, y[i].im);
                }
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    r[i] = r[i] + dt6 * (ar[i] + vs[i]);
                    v[i] = v[i] + dt6 * (av[i] + F[i] / m);
                }
            }
        }
        if (stop) break;
        t += dt;
        // jump check with the norm in double
        double ns = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) ns += (double)y[i].re * (double)y[i].re + (double)y[i].im * (double)y[i].im;
        if (1.0 - ns > thr) {
            n_jump++;
            const mm_u32x4 wa = mc_draw(R.seed, tid, 3u * n_jump - 2u, 0u);
            double pe[NE], tot = 0.0;
#pragma unroll
            for (int e = 0; e < NE; e++) {
                pe[e] = (double)y[NG + e].re * (double)y[NG + e].re + (double)y[NG + e].im * (double)y[NG + e].im;
                if (pe[e] < 1e-30) continue;
                for (int g = 0; g < NG; g++)
                    for (int q = 0; q < 3; q++) {
                        const double b = P->br[e][g][q];
                        if (b < 1e-15) continue;
                        tot += b * pe[e];
                    }
            }
            int sel_g = 0, sel_ch = -1;
            if (tot > 1e-30) {
                const double rc = mc_uniform(wa.x, wa.y) * tot;
                double cum = 0.0;
                for (int e = 0; e < NE && sel_ch < 0; e++) {
                    if (pe[e] < 1e-30) continue;
                    for (int g = 0; g < NG && sel_ch < 0; g++)
                        for (int q = 0; q < 3 && sel_ch < 0; q++) {
                            const double b = P->br[e][g][q];
                            if (b < 1e-15) continue;
                            cum += b * pe[e];
                            if (rc <= cum) {
                                sel_g = g;
                                sel_ch = (e * NG + g) * 3 + q;
                            }
                        }
                }
                if (sel_ch < 0) sel_ch = 0;  // unreachable: rc <= tot
                n_sc++;
                if (R.jump_cap > 0 && n_logged < R.jump_cap) {
                    R.jump_t[tid * R.jump_cap + n_logged] = t;
                    R.jump_ch[tid * R.jump_cap + n_logged] = sel_ch;
                    n_logged++;
                }
                if (P->add_kick) {
                    const mm_u32x4 wb = mc_draw(R.seed, tid, 3u * n_jump - 1u, 0u);
                    const double ct = 2.0 * mc_uniform(wb.x, wb.y) - 1.0;
                    const double st = sqrt(1.0 - ct * ct);
                    const double phk = 2.0 * M_PI * mc_uniform(wb.z, wb.w);
                    v[0] += P->v_recoil * (st * cos(phk));
                    v[1] += P->v_recoil * (st * sin(phk));
                    v[2] += P->v_recoil * ct;
                }
            }
#pragma unroll
            for (int i = 0; i < N; i++) {
                y[i].re = (i == sel_g) ? T(1) : T(0);
                y[i].im = T(0);
            }
            if (P->diffusion[0] > 0.0 || P->diffusion[1] > 0.0 || P->diffusion[2] > 0.0) {
                const mm_u32x4 wd = mc_draw(R.seed, tid, 3u * n_jump, 0u);
                const uint32_t wds[3] = {wd.x, wd.y, wd.z};
                const double dn = t - last;
#pragma unroll
                for (int i = 0; i < 3; i++)
                    if (P->diffusion[i] > 0.0)
                        v[i] += (wds[i] >= 0x80000000u ? 1.0 : -1.0) * sqrt(2.0 * P->diffusion[i] * dn);
            }
            last = t;
            thr = mc_threshold(mc_uniform(wa.z, wa.w), R.jump_rule);
        }
        if (R.renorm_interval > 0 && (step + 1) % R.renorm_interval == 0) {
            double nr = 0.0;
#pragma unroll
            for (int i = 0; i < N; i++)
                nr += (double)y[i].re * (double)y[i].re + (double)y[i].im * (double)y[i].im;
            if (nr > 1.0 + 1e-10 || nr < 1e-15) {
                const T s = (T)(1.0 / sqrt(nr));
#pragma unroll
                for (int i = 0; i < N; i++) {
                    y[i].re *= s;
                    y[i].im *= s;
                }
                n_ren++;
                const mm_u32x4 wr = mc_draw(R.seed, tid, n_ren, 1u);
                thr = mc_threshold(mc_uniform(wr.x, wr.y), R.jump_rule);
            }
        }
    }
    R.n_scatters[tid] = n_sc;
    R.escaped[tid] = esc;
    R.n_saved[tid] = si;
    if (R.n_avg) R.n_avg[tid] = n_av;
    if (R.fsum) {
#pragma unroll
        for (int i = 0; i < 3; i++) R.fsum[tid * 3 + i] = fs[i];
    }
    if (R.r_final) {
#pragma unroll
        for (int i = 0; i < 3; i++) R.r_final[tid * 3 + i] = r[i];
    }
    if (R.v_final) {
#pragma unroll
        for (int i = 0; i < 3; i++) R.v_final[tid * 3 + i] = v[i];
    }
    if (R.psi_final) {
#pragma unroll
        for (int i = 0; i < N; i++) {
            R.psi_final[(tid * N + i) * 2] = (double)y[i].re;
            R.psi_final[(tid * N + i) * 2 + 1] = (double)y[i].im;
        }
    }
}

// Laser sums for every stage time: S[tau][b][q] = sum_f amp_{b,f,q} exp(i w_f t_tau).
template <typename T>
__global__ void k_mcwf_fields(const McProblem* __restrict__ P, long n_tau, const double* __restrict__ t_tau,
                              C2<T>* __restrict__ S) {
    const long tau = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tau >= n_tau) return;
    const double t = t_tau[tau];
    double cs[MC_MAXF], sn[MC_MAXF];
    for (int f = 0; f < P->nf; f++) sincos(P->wl[f] * t, &sn[f], &cs[f]);
    for (int b = 0; b < 6; b++)
        for (int q = 0; q < 3; q++) {
            double re = 0.0, imv = 0.0;
            for (int f = 0; f < P->nf; f++) {
                re += P->amp_re[b][f][q] * cs[f] - P->amp_im[b][f][q] * sn[f];
                imv += P->amp_re[b][f][q] * sn[f] + P->amp_im[b][f][q] * cs[f];
            }
            S[tau * 18 + b * 3 + q].re = (T)re;
            S[tau * 18 + b * 3 + q].im = (T)imv;
        }
}

#define MC_SIZE_LIST(X) X(12, 4) X(1, 1) X(2, 2)

extern "C" {

int mm_mcwf_problem_create(mm_mcwf_problem** out, int ng, int ne, int nf, const double* wbase,
                           const double* zeeman_re, const double* zeeman_im, const double* zdiag_e,
                           const double* d, const double* br, const double* amp_re, const double* amp_im,
                           const double* w_lasers, double B_gradient, double r_nat, double beam_radius_k,
                           double mass_nat, double v_recoil, int add_kick, const double* diffusion) {
    mm_clear_error();
    if (!out) return mm_fail(MM_EINVAL, "mcwf_problem_create: null output");
    *out = nullptr;
    if (!wbase || !zeeman_re || !zeeman_im || !zdiag_e || !d || !br || !amp_re || !amp_im || !w_lasers ||
        !diffusion)
        return mm_fail(MM_EINVAL, "mcwf_problem_create: null table");
    bool ok = false;
#define X(G, E) ok = ok || (ng == G && ne == E);
    MC_SIZE_LIST(X)
#undef X
    if (!ok) return mm_fail(MM_EUNSUPPORTED, "mcwf: no kernel for %d ground + %d excited states", ng, ne);
    if (nf < 1 || nf > MC_MAXF) return mm_fail(MM_EINVAL, "mcwf: %d laser frequencies (1..%d)", nf, MC_MAXF);
    if (!(mass_nat > 0.0) || !(beam_radius_k > 0.0)) return mm_fail(MM_EINVAL, "mcwf: mass and beam radius > 0");
    mm_mcwf_problem* p = new mm_mcwf_problem();
    McProblem& h = p->host;
    h.ng = ng;
    h.ne = ne;
    h.nf = nf;
    h.add_kick = add_kick ? 1 : 0;
    for (int n = 0; n < ng + ne; n++) h.wbase[n] = wbase[n];
    for (int e = 0; e < ne; e++) h.zde[e] = zdiag_e[e];
    for (int a = 0; a < 3; a++)
        for (int i = 0; i < ng; i++) {
            h.zgd[i][a] = zeeman_re[(a * ng + i) * ng + i];
            for (int j = 0; j < ng; j++) {
                const double zr = zeeman_re[(a * ng + i) * ng + j], zi = zeeman_im[(a * ng + i) * ng + j];
                h.td.zr[i][j][a] = zr;
                h.td.zi[i][j][a] = zi;
                h.tf.zr[i][j][a] = (float)zr;
                h.tf.zi[i][j][a] = (float)zi;
            }
        }
    for (int g = 0; g < ng; g++)
        for (int e = 0; e < ne; e++)
            for (int q = 0; q < 3; q++) {
                h.td.d[g][e][q] = d[(g * ne + e) * 3 + q];
                h.tf.d[g][e][q] = (float)d[(g * ne + e) * 3 + q];
                h.br[e][g][q] = br[(e * ng + g) * 3 + q];
            }
    for (int b = 0; b < 6; b++)
        for (int f = 0; f < nf; f++)
            for (int q = 0; q < 3; q++) {
                h.amp_re[b][f][q] = amp_re[(b * nf + f) * 3 + q];
                h.amp_im[b][f][q] = amp_im[(b * nf + f) * 3 + q];
            }
    for (int f = 0; f < nf; f++) h.wl[f] = w_lasers[f];
    h.B_gradient = B_gradient;
    h.r_nat = r_nat;
    h.denom = beam_radius_k * beam_radius_k / 2.0;
    h.mass_nat = mass_nat;
    h.v_recoil = v_recoil;
    for (int i = 0; i < 3; i++) h.diffusion[i] = diffusion[i];
    if (hipMalloc(&p->dev, sizeof(McProblem)) != hipSuccess) {
        delete p;
        return mm_hip_fail("mcwf_problem_create: hipMalloc");
    }
    if (hipMemcpy(p->dev, &h, sizeof(McProblem), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p->dev);
        delete p;
        return mm_hip_fail("mcwf_problem_create: hipMemcpy");
    }
    *out = p;
    return MM_OK;
}

int mm_mcwf_problem_destroy(mm_mcwf_problem* p) {
    mm_clear_error();
    if (!p) return MM_OK;
    (void)hipFree(p->dev);
    delete p;
    return MM_OK;
}

int mm_mcwf_field_table(const mm_mcwf_problem* p, long n_tau, const double* t_tau, int precision, void* S,
                        void* stream) {
    mm_clear_error();
    if (!p) return mm_fail(MM_EINVAL, "mcwf_field_table: null problem");
    if (n_tau < 0) return mm_fail(MM_EINVAL, "mcwf_field_table: negative size");
    if (n_tau == 0) return MM_OK;
    if (!t_tau || !S) return mm_fail(MM_EINVAL, "mcwf_field_table: null pointer");
    const unsigned grid = (unsigned)((n_tau + 255) / 256);
    if (precision == 0)
        hipLaunchKernelGGL(k_mcwf_fields<double>, dim3(grid), dim3(256), 0, (hipStream_t)stream, p->dev, n_tau, t_tau,
                           (C2<double>*)S);
    else if (precision == 1)
        hipLaunchKernelGGL(k_mcwf_fields<float>, dim3(grid), dim3(256), 0, (hipStream_t)stream, p->dev, n_tau, t_tau,
                           (C2<float>*)S);
    else
        return mm_fail(MM_EINVAL, "mcwf_field_table: precision %d (0 fp64, 1 fp32)", precision);
    return mm_check_launch("k_mcwf_fields");
}

int mm_mcwf_run(const mm_mcwf_problem* p, long n, const double* r0, const double* v0, const double* psi0,
                const void* field_table, long n_tau, const mm_mcwf_config* cfg, const mm_mcwf_outputs* out,
                void* stream) {
    mm_clear_error();
    if (!p || !cfg || !out) return mm_fail(MM_EINVAL, "mcwf_run: null problem, config or outputs");
    if (n < 0) return mm_fail(MM_EINVAL, "mcwf_run: negative size");
    if (cfg->n_steps < 0 || cfg->save_every < 1 || cfg->n_save < 0)
        return mm_fail(MM_EINVAL, "mcwf_run: n_steps >= 0, save_every >= 1, n_save >= 0");
    if (!(cfg->dt > 0.0)) return mm_fail(MM_EINVAL, "mcwf_run: dt must be positive");
    if (cfg->n_steps > (1 << 29)) return mm_fail(MM_EINVAL, "mcwf_run: n_steps too large");
    if (cfg->jump_rule != 0 && cfg->jump_rule != 1) return mm_fail(MM_EINVAL, "mcwf_run: jump_rule 0 or 1");
    if (cfg->jump_cap < 0 || cfg->renorm_interval < 0)
        return mm_fail(MM_EINVAL, "mcwf_run: negative jump_cap or renorm_interval");
    if (n == 0) return MM_OK;
    if (!r0 || !v0 || !field_table || !out->n_scatters || !out->escaped || !out->n_saved)
        return mm_fail(MM_EINVAL, "mcwf_run: null pointer");
    if (cfg->jump_cap > 0 && (!out->jump_t || !out->jump_ch))
        return mm_fail(MM_EINVAL, "mcwf_run: jump log buffers missing");
    if (n_tau < 2L * cfg->n_steps + 2)
        return mm_fail(MM_EINVAL, "mcwf_run: field table has %ld entries, needs %ld", n_tau,
                       2L * cfg->n_steps + 2);
    McRun R;
    R.n = n;
    R.dt = cfg->dt;
    R.r_escape2 = cfg->r_escape > 0.0 ? cfg->r_escape * cfg->r_escape : 0.0;
    R.n_steps = cfg->n_steps;
    R.save_every = cfg->save_every;
    R.n_save = cfg->n_save;
    R.renorm_interval = cfg->renorm_interval;
    R.avg_start = cfg->avg_start;
    R.jump_cap = cfg->jump_cap;
    R.jump_rule = cfg->jump_rule;
    R.seed = cfg->seed;
    R.r0 = r0;
    R.v0 = v0;
    R.psi0 = psi0;
    R.S = field_table;
    R.r_out = out->r;
    R.v_out = out->v;
    R.pop_out = out->pop;
    R.f_out = out->force;
    R.fsum = out->force_sum;
    R.psi_final = out->psi_final;
    R.r_final = out->r_final;
    R.v_final = out->v_final;
    R.jump_t = out->jump_t;
    R.n_scatters = out->n_scatters;
    R.escaped = out->escaped;
    R.n_saved = out->n_saved;
    R.n_avg = out->n_avg;
    R.jump_ch = out->jump_ch;
    const unsigned grid = (unsigned)((n + 63) / 64);
    const int ng = p->host.ng, ne = p->host.ne;
#define X(G, E)                                                                                                \
    if (ng == G && ne == E) {                                                                                  \
        if (cfg->precision == 0)                                                                               \
            hipLaunchKernelGGL((k_mcwf<double, G, E>), dim3(grid), dim3(64), 0, (hipStream_t)stream, p->dev, R); \
        else if (cfg->precision == 1)                                                                          \
            hipLaunchKernelGGL((k_mcwf<float, G, E>), dim3(grid), dim3(64), 0, (hipStream_t)stream, p->dev, R);  \
        else                                                                                                   \
            return mm_fail(MM_EINVAL, "mcwf_run: precision %d (0 fp64, 1 fp32)", cfg->precision);             \
        return mm_check_launch("k_mcwf");                                                                      \
    }
    MC_SIZE_LIST(X)
#undef X
    return mm_fail(MM_EUNSUPPORTED, "mcwf_run: no kernel for %d + %d", ng, ne);
}

}  // extern "C"
