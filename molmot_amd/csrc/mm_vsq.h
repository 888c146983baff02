// Monte-Carlo temperature statistics of the 1D ensemble
// (molmot/propagation/trajectories.py:159-229): per particle, the sum of v^2
// and the number of samples over steps j >= i_eq that end inside
// |z| <= z_escape.  The bodies of the two launch forms (one thread per
// particle; the work queue of mm_queue_dev.h) over an evaluator EV of the
// force (mm_traj.h's interface), so the kernels with the tables in memory
// (traj3d.hip) and the plan-specialised ones (rate_spec.hip) share the step.
// Free of host headers: run-time compiled kernels include it.
#pragma once

#include "mm_queue_dev.h"
#include "mm_traj.h"

// Steps j0 <= j < j1 from (z, v) with the sums (acc, c); stops after the
// step that leaves |z| > z_escape (its sample is not taken).  RF: the force
// is the mean of P and of P2 at -Bgrad (the two switching phases of the RF
// MOT, molmot/mot/simulator.py:49-90), P2 evaluated by EV2.
template <class EV, int NG, int NE, bool UNIT, int BS, bool RF, class EV2 = MMEvalPlan<NG, NE, UNIT, 1, false>>
__device__ __forceinline__ void vsq_steps(const MMRatePlanDev* __restrict__ P, const MMRatePlanDev* __restrict__ P2,
                                          int j0, int j1, double& z, double& v, double& acc, long long& c,
                                          double Bgrad, double mass, double inv_mass, double dt, int i_eq,
                                          double z_escape, double* lds) {
    for (int j = j0; j < j1; j++) {
        double f, rr;
        EV::eval(P, v, z, Bgrad, f, rr, lds, BS);
        if (RF) {
            double f2;
            EV2::eval(P2, v, z, -Bgrad, f2, rr, lds, BS);
            f = 0.5 * (f + f2);
        }
        const double a = mm_div_rn(f, mass, inv_mass);
        v = v + a * dt;
        z = z + v * dt;
        if (fabs(z) > z_escape) break;
        if (j >= i_eq) {
            acc = acc + v * v;
            c += 1;
        }
    }
}

// One thread per particle (k_traj_vsq's contract).
template <class EV, int NG, int NE, bool UNIT, int BS, bool RF, class EV2 = MMEvalPlan<NG, NE, UNIT, 1, false>>
__device__ __forceinline__ void vsq_grid_body(const MMRatePlanDev* __restrict__ P,
                                              const MMRatePlanDev* __restrict__ P2, long i, long n,
                                              const double* __restrict__ z0, const double* __restrict__ v0,
                                              double Bgrad, double mass, double dt, int n_steps, int i_eq,
                                              double z_escape, double* __restrict__ vsq,
                                              long long* __restrict__ cnt, double* lds) {
    if (i >= n) return;
    const double inv_mass = 1.0 / mass;
    double z = z0[i], v = v0[i], acc = 0.0;
    long long c = 0;
    vsq_steps<EV, NG, NE, UNIT, BS, RF, EV2>(P, P2, 1, n_steps, z, v, acc, c, Bgrad, mass, inv_mass, dt, i_eq,
                                             z_escape, lds);
    vsq[i] = acc;
    cnt[i] = c;
}

// Work-queue form (k_traj_vsq_queue's contract): a particle's (z, v) live in
// zs, vs and its sums in vsq, cnt between chunks of steps; it continues
// while it is inside z_escape (the check follows every step, so a stored
// state outside means it escaped).
template <class EV, int NG, int NE, bool UNIT>
__device__ __forceinline__ void vsq_queue_body(const MMRatePlanDev* __restrict__ P, long n,
                                               const double* __restrict__ z0, const double* __restrict__ v0,
                                               double Bgrad, double mass, double dt, int n_steps, int i_eq,
                                               double z_escape, double* __restrict__ vsq,
                                               long long* __restrict__ cnt, double* __restrict__ zs,
                                               double* __restrict__ vs, const MMQueue& Q, double* lds) {
    const double inv_mass = 1.0 / mass;
    unsigned g, ch;
    while (mm_queue_next(Q, g, ch)) {
        const long i = (long)g * 64 + threadIdx.x;
        if (i < n) {
            const int j0 = 1 + (int)ch * Q.chunk;
            const int j1 = min(n_steps, j0 + Q.chunk);
            double z, v, acc;
            long long c;
            if (ch == 0) {
                z = z0[i];
                v = v0[i];
                acc = 0.0;
                c = 0;
            } else {
                z = zs[i];
                v = vs[i];
                acc = vsq[i];
                c = cnt[i];
            }
            if (ch == 0 || !(fabs(z) > z_escape))
                vsq_steps<EV, NG, NE, UNIT, 64, false>(P, nullptr, j0, j1, z, v, acc, c, Bgrad, mass, inv_mass, dt,
                                                       i_eq, z_escape, lds);
            zs[i] = z;
            vs[i] = v;
            vsq[i] = acc;
            cnt[i] = c;
        }
        mm_queue_publish(Q, g, ch);
    }
}
