// Symplectic-Euler step loop of the 1D ensemble (molmot/propagation/
// trajectories_jit.py:18-58) and the bodies of its two launch forms (one
// thread per particle; the work queue of mm_queue_dev.h), written once over an
// evaluator EV of the rate-equation force:
//     EV::eval(P, v, z, Bgrad, F, R, lds, ls)
// MMEvalPlan reads the plan's tables from device memory (rate_core.h); the
// plan-specialised evaluators that rate_spec.hip generates at run time hold
// the tables as constants of their own code.  Every evaluator returns the
// same F and R bits for the same plan, so every kernel built on this file
// integrates the same trajectories.
// Free of host headers: run-time compiled kernels include it too.
#pragma once

#include "mm_plan.h"
#include "mm_queue_dev.h"
#include "philox.h"
#include "rate_core.h"

// Ground-state loops of the generic trajectory kernels unrolled by 6: fewer
// loop and address instructions and two ground states' pair terms in
// flight; the register count (~150) still allows more than the two waves per
// SIMD an ensemble of 1e5 particles provides.
#ifndef MM_TRAJ_GU
#define MM_TRAJ_GU 6
#endif

// The tables in device memory.  PK: the variant for packed plans (one pair
// path in the loop: +9% at 1e5 particles).
template <int NG, int NE, bool UNIT, int GU, bool PK>
struct MMEvalPlan {
    __device__ __forceinline__ static void eval(const MMRatePlanDev* __restrict__ P, double v, double z, double Bgrad,
                                                double& F, double& R, double* lds, int ls) {
        rate_eval<NG, NE, UNIT, GU, PK>(P, v, z, Bgrad, F, R, nullptr, lds, ls);
    }
};

// Euler steps j0 <= j < j1 of particle i from (z, v); na becomes the escape
// step if the particle leaves.  j0 is odd when KICKS: a Philox block serves
// an odd step and the even step after it, so a run of steps may start at any
// odd step with the same normals as an unbroken one.  RF: the force (and
// scattering rate) is the mean of two plans, the second evaluated at -Bgrad
// -- the two switching phases of the RF MOT (molmot/mot/simulator.py:49-90);
// P2 is unused otherwise.  EV2 evaluates the second phase.
template <class EV, int NG, int NE, bool UNIT, bool SAVE, bool KICKS, int BS, bool RF,
          class EV2 = MMEvalPlan<NG, NE, UNIT, 1, false>>
__device__ __forceinline__ void traj_steps_ev(const MMRatePlanDev* __restrict__ P,
                                              const MMRatePlanDev* __restrict__ P2, long i, int j0, int j1,
                                              int n_steps, double& z, double& v, int& na, double Bgrad, double mass,
                                              double inv_mass, double dt, double z_escape, double* __restrict__ zt,
                                              double* __restrict__ vt, double kick_scale, uint32_t k0, uint32_t k1,
                                              uint32_t sid, double* lds, float* xi_carry = nullptr) {
    // xi_carry (segmented runs that may stop after an odd step): the sine
    // normal of the last Philox pair, handed to the next segment
    float xi_now = 0.f, xi_next = xi_carry ? *xi_carry : 0.f;
    for (int j = j0; j < j1; j++) {
        double f, r;
        EV::eval(P, v, z, Bgrad, f, r, lds, BS);
        if (RF) {
            double f2, r2;
            EV2::eval(P2, v, z, -Bgrad, f2, r2, lds, BS);
            f = 0.5 * (f + f2);
            r = 0.5 * (r + r2);
        }
        const double a = mm_div_rn(f, mass, inv_mass);
        v = v + a * dt;
        if (KICKS) {
            // one Philox block and one Box-Muller pair per two steps: the
            // cosine normal kicks odd step j, the sine normal step j + 1
            // (counter = (j >> 1, particle, stream))
            if (j & 1) {
                const mm_u32x4 c = {(uint32_t)(j >> 1), (uint32_t)i, sid, (uint32_t)(i >> 32)};
                const mm_u32x4 o = mm_philox4x32_10(c, k0, k1);
                mm_normal2_from(o.x, o.y, xi_now, xi_next);
            } else {
                xi_now = xi_next;
            }
            // kick size in FP32 (a random kick; 1e-7 relative is immaterial)
            v = fma(kick_scale, (double)(sqrtf((float)(r * dt)) * xi_now), v);
        }
        z = z + v * dt;
        if (SAVE) {
            zt[i * n_steps + j] = z;
            vt[i * n_steps + j] = v;
        }
        if (fabs(z) > z_escape) {
            na = j;
            if (SAVE) {
                for (int s = j + 1; s < n_steps; s++) {
                    zt[i * n_steps + s] = z;
                    vt[i * n_steps + s] = v;
                }
            }
            break;
        }
    }
    if (xi_carry) *xi_carry = xi_next;
}

// The generic kernels' step loop (tables in device memory).
template <int NG, int NE, bool UNIT, bool SAVE, bool KICKS, int BS, bool RF, bool PK>
__device__ __forceinline__ void traj_steps(const MMRatePlanDev* __restrict__ P, const MMRatePlanDev* __restrict__ P2,
                                           long i, int j0, int j1, int n_steps, double& z, double& v, int& na,
                                           double Bgrad, double mass, double inv_mass, double dt, double z_escape,
                                           double* __restrict__ zt, double* __restrict__ vt, double kick_scale,
                                           uint32_t k0, uint32_t k1, uint32_t sid, double* lds,
                                           float* xi_carry = nullptr) {
    traj_steps_ev<MMEvalPlan<NG, NE, UNIT, MM_TRAJ_GU, PK>, NG, NE, UNIT, SAVE, KICKS, BS, RF>(
        P, P2, i, j0, j1, n_steps, z, v, na, Bgrad, mass, inv_mass, dt, z_escape, zt, vt, kick_scale, k0, k1, sid,
        lds, xi_carry);
}

// One thread integrates particle i for all steps; escaped particles leave
// the loop (their lanes idle until the wave's other particles finish).
template <class EV, int NG, int NE, bool UNIT, bool SAVE, bool KICKS, int BS, bool RF,
          class EV2 = MMEvalPlan<NG, NE, UNIT, 1, false>>
__device__ __forceinline__ void traj_grid_body(const MMRatePlanDev* __restrict__ P, long i, long n,
                                               const double* __restrict__ z0, const double* __restrict__ v0,
                                               double Bgrad, double mass, double dt, int n_steps, double z_escape,
                                               double* __restrict__ zt, double* __restrict__ vt,
                                               double* __restrict__ zf, double* __restrict__ vf,
                                               int* __restrict__ nact, double kick_scale, uint32_t k0, uint32_t k1,
                                               uint32_t sid, const MMRatePlanDev* __restrict__ P2, double* lds) {
    if (i >= n) return;
    const double inv_mass = 1.0 / mass;
    double z = z0[i], v = v0[i];
    int na = n_steps;
    if (SAVE) {
        zt[i * n_steps] = z;
        vt[i * n_steps] = v;
    }
    traj_steps_ev<EV, NG, NE, UNIT, SAVE, KICKS, BS, RF, EV2>(P, P2, i, 1, n_steps, n_steps, z, v, na, Bgrad,
                                                              mass, inv_mass, dt, z_escape, zt, vt, kick_scale, k0,
                                                              k1, sid, lds);
    zf[i] = z;
    vf[i] = v;
    nact[i] = na;
}

// Persistent form (no saved trajectory, no RF): wavefronts take (group of 64
// particles, chunk of steps) units from the queue of mm_queue_dev.h.  A
// particle's (z, v, n_actual) live in the output arrays between chunks; it
// continues while n_actual = n_steps (not escaped).  Chunks start at odd
// steps, so the Philox pairs and every step's arithmetic are those of the
// one-thread-per-particle form: the results are bit-identical.
template <class EV, int NG, int NE, bool UNIT, bool KICKS>
__device__ __forceinline__ void traj_queue_body(const MMRatePlanDev* __restrict__ P, long n,
                                                const double* __restrict__ z0, const double* __restrict__ v0,
                                                double Bgrad, double mass, double dt, int n_steps, double z_escape,
                                                double* __restrict__ zf, double* __restrict__ vf,
                                                int* __restrict__ nact, double kick_scale, uint32_t k0, uint32_t k1,
                                                uint32_t sid, const MMQueue& Q, double* lds) {
    const double inv_mass = 1.0 / mass;
    unsigned g, c;
    while (mm_queue_next(Q, g, c)) {
        const long i = (long)g * 64 + threadIdx.x;
        const int j0 = 1 + (int)c * Q.chunk;
        const int j1 = min(n_steps, j0 + Q.chunk);
        if (i < n) {
            double z, v;
            int na;
            if (c == 0) {
                z = z0[i];
                v = v0[i];
                na = n_steps;
            } else {
                z = zf[i];
                v = vf[i];
                na = nact[i];
            }
            if (na == n_steps)
                traj_steps_ev<EV, NG, NE, UNIT, false, KICKS, 64, false>(P, nullptr, i, j0, j1, n_steps, z, v, na,
                                                                         Bgrad, mass, inv_mass, dt, z_escape,
                                                                         nullptr, nullptr, kick_scale, k0, k1, sid,
                                                                         lds);
            zf[i] = z;
            vf[i] = v;
            nact[i] = na;
        }
        mm_queue_publish(Q, g, c);
    }
}

// ---------------------------------------------------------------------------
// Ensemble statistics at sample steps (molmot/mot/analysis.py:128-288 reads
// z and v of every trajectory at the samples nearest the time-bin centres,
// keeping |z| < r_alive): each particle runs the ensemble's Euler steps
// segment by segment (bit-identical to one unbroken run, kicks included)
// and, at every sample step, contributes (1, z, v) if alive.  Per sample the
// wavefront merges (count, mean, M2) of z and v pairwise (Chan et al.), one
// partial per 64-thread block goes to `part`, and a second kernel merges the
// block partials in block order: deterministic, and no one-pass
// cancellation in the variance.
// ---------------------------------------------------------------------------
struct MMMoments {
    double n, mz, m2z, mv, m2v;
};

__device__ __forceinline__ void mm_moments_merge(MMMoments& a, const MMMoments& b) {
    const double n = a.n + b.n;
    if (b.n == 0.0) return;
    if (a.n == 0.0) {
        a = b;
        return;
    }
    const double wb = b.n / n;
    const double dz = b.mz - a.mz, dv = b.mv - a.mv;
    a.mz = fma(dz, wb, a.mz);
    a.mv = fma(dv, wb, a.mv);
    a.m2z = a.m2z + b.m2z + dz * dz * a.n * wb;
    a.m2v = a.m2v + b.m2v + dv * dv * a.n * wb;
    a.n = n;
}

// One thread per particle, 64-thread blocks; samp[0..n_samp) are the state
// indices sampled (0 <= samp[k] < n_steps, non-decreasing); block b writes
// part[(b * n_samp + k) * 5 .. + 5) = (count, mean z, M2 z, mean v, M2 v).
template <class EV, int NG, int NE, bool UNIT, bool KICKS>
__device__ __forceinline__ void traj_stats_body(const MMRatePlanDev* __restrict__ P, long n,
                                                const double* __restrict__ z0, const double* __restrict__ v0,
                                                double Bgrad, double mass, double dt, int n_steps, double z_escape,
                                                double r_alive, const int* __restrict__ samp, int n_samp,
                                                double* __restrict__ part, double kick_scale, uint32_t k0,
                                                uint32_t k1, uint32_t sid, double* lds) {
    const long i = (long)blockIdx.x * 64 + threadIdx.x;
    const bool live = i < n;
    const double inv_mass = 1.0 / mass;
    double z = live ? z0[i] : 0.0, v = live ? v0[i] : 0.0;
    int na = n_steps;
    float xi = 0.f;
    int cur = 1;  // next step to integrate (state index cur - 1 is current)
    for (int k = 0; k < n_samp; k++) {
        const int target = samp[k];
        if (live && na == n_steps && target >= cur) {
            traj_steps_ev<EV, NG, NE, UNIT, false, KICKS, 64, false>(P, nullptr, i, cur, target + 1, n_steps, z, v,
                                                                     na, Bgrad, mass, inv_mass, dt, z_escape,
                                                                     nullptr, nullptr, kick_scale, k0, k1, sid, lds,
                                                                     &xi);
        }
        if (target >= cur) cur = target + 1;
        MMMoments m = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (live && fabs(z) < r_alive) m = {1.0, z, 0.0, v, 0.0};
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            MMMoments b;
            b.n = __shfl_xor(m.n, o);
            b.mz = __shfl_xor(m.mz, o);
            b.m2z = __shfl_xor(m.m2z, o);
            b.mv = __shfl_xor(m.mv, o);
            b.m2v = __shfl_xor(m.m2v, o);
            // merge in lane order so that both partners hold the same value
            if (threadIdx.x & o) {
                mm_moments_merge(b, m);
                m = b;
            } else {
                mm_moments_merge(m, b);
            }
        }
        if (threadIdx.x == 0) {
            double* q = part + ((long)blockIdx.x * n_samp + k) * 5;
            q[0] = m.n;
            q[1] = m.mz;
            q[2] = m.m2z;
            q[3] = m.mv;
            q[4] = m.m2v;
        }
    }
}
