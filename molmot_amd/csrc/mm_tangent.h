// Forward sensitivities of the 1D ensemble over an evaluator GE of the force
// gradient (the plan-specialised kernels of rate_spec.hip):
//     GE::grad(v, z, Bgrad, F, J) -> bool  (false: the reduced solve does not apply)
// The state (z, v) and its 2 x 9 sensitivities S_z = dz/dp, S_v = dv/dp
// (p = 4 detunings, split, s0 scale, B', z_0, v_0) advance together:
//     S_v <- S_v + dt/m (dF/dz S_z + dF/dv S_v + dF/dp),  S_z <- S_z + dt S_v
// with the update of k_traj_tangent / k_traj_tangent_queue (rate_adjoint.hip)
// operation for operation, so that only F and J can differ between the two.
// Free of host headers: run-time compiled kernels include it.
#pragma once

#include "mm_queue_dev.h"
#include "rate_adjoint.h"

#ifndef MM_TAN_NP
#define MM_TAN_NP (MM_ADJ_NJ)
#endif

// Updates k < k_end of one particle; stops after the update that leaves
// |z| > z_escape.  Where GE declines, F comes from the dense solve and J is
// NaN (as in the generic kernel).
template <class GE, int NG, int NE>
__device__ __forceinline__ void tangent_steps_ge(const MMRatePlanDev* __restrict__ P, double& z, double& v, int& k,
                                                 int k_end, double (&Sz)[MM_TAN_NP], double (&Sv)[MM_TAN_NP],
                                                 double Bgrad, double mass, double inv_mass, double dt,
                                                 double z_escape) {
    const double c = dt / mass;
    while (k < k_end) {
        double f, J[MM_ADJ_NJ];
        if (!GE::grad(v, z, Bgrad, f, J)) {
            double fd, r;
            rate_eval_dense<NG, NE>(P, v, z, Bgrad, &fd, &r, nullptr);
            f = fd;
#pragma unroll
            for (int j = 0; j < MM_ADJ_NJ; j++) J[j] = __builtin_nan("");
        }
        const double Jz = J[MM_ADJ_J_Z], Jv = J[MM_ADJ_J_V];
#pragma unroll
        for (int j = 0; j < MM_TAN_NP; j++) {
            double dF = fma(Jz, Sz[j], Jv * Sv[j]);
            if (j < MM_ADJ_NJ - 2) dF += J[j];
            Sv[j] = fma(c, dF, Sv[j]);
            Sz[j] = fma(dt, Sv[j], Sz[j]);
        }
        const double a = mm_div_rn(f, mass, inv_mass);
        v = v + a * dt;
        z = z + v * dt;
        k++;
        if (fabs(z) > z_escape) break;
    }
}

// One thread per particle (k_traj_tangent's contract: n_up[i] >= 0 limits
// particle i's updates; sens[i] = (S_z[0..8], S_v[0..8])).
template <class GE, int NG, int NE>
__device__ __forceinline__ void tangent_grid_body(const MMRatePlanDev* __restrict__ P, long i, long n,
                                                  const double* __restrict__ z0, const double* __restrict__ v0,
                                                  double Bgrad, double mass, double dt, int n_steps,
                                                  const int* __restrict__ n_up, double z_escape,
                                                  double* __restrict__ zf, double* __restrict__ vf,
                                                  int* __restrict__ n_done, double* __restrict__ sens) {
    if (i >= n) return;
    const double inv_mass = 1.0 / mass;
    double z = z0[i], v = v0[i];
    int lim = n_steps - 1;
    if (n_up && n_up[i] >= 0 && n_up[i] < lim) lim = n_up[i];
    double Sz[MM_TAN_NP], Sv[MM_TAN_NP];
#pragma unroll
    for (int j = 0; j < MM_TAN_NP; j++) {
        Sz[j] = (j == MM_ADJ_NJ - 2) ? 1.0 : 0.0;
        Sv[j] = (j == MM_ADJ_NJ - 1) ? 1.0 : 0.0;
    }
    int k = 0;
    tangent_steps_ge<GE, NG, NE>(P, z, v, k, lim, Sz, Sv, Bgrad, mass, inv_mass, dt, z_escape);
    zf[i] = z;
    vf[i] = v;
    n_done[i] = k;
#pragma unroll
    for (int j = 0; j < MM_TAN_NP; j++) {
        sens[i * 2 * MM_TAN_NP + j] = Sz[j];
        sens[i * 2 * MM_TAN_NP + MM_TAN_NP + j] = Sv[j];
    }
}

// Work-queue form (k_traj_tangent_queue's contract): a particle's (z, v,
// updates done, sensitivities) live in the outputs between chunks; it
// continues while it has run every earlier update and is inside z_escape.
template <class GE, int NG, int NE>
__device__ __forceinline__ void tangent_queue_body(const MMRatePlanDev* __restrict__ P, long n,
                                                   const double* __restrict__ z0, const double* __restrict__ v0,
                                                   double Bgrad, double mass, double dt, int n_steps,
                                                   const int* __restrict__ n_up, double z_escape,
                                                   double* __restrict__ zf, double* __restrict__ vf,
                                                   int* __restrict__ n_done, double* __restrict__ sens,
                                                   const MMQueue& Q) {
    const double inv_mass = 1.0 / mass;
    unsigned g, ch;
    while (mm_queue_next(Q, g, ch)) {
        const long i = (long)g * 64 + threadIdx.x;
        if (i < n) {
            int lim = n_steps - 1;
            if (n_up && n_up[i] >= 0 && n_up[i] < lim) lim = n_up[i];
            const int k0 = (int)ch * Q.chunk;
            const int k1 = min(lim, k0 + Q.chunk);
            double z, v, Sz[MM_TAN_NP], Sv[MM_TAN_NP];
            int k;
            double* si = sens + i * 2 * MM_TAN_NP;
            if (ch == 0) {
                z = z0[i];
                v = v0[i];
                k = 0;
#pragma unroll
                for (int j = 0; j < MM_TAN_NP; j++) {
                    Sz[j] = (j == MM_ADJ_NJ - 2) ? 1.0 : 0.0;
                    Sv[j] = (j == MM_ADJ_NJ - 1) ? 1.0 : 0.0;
                }
            } else {
                z = zf[i];
                v = vf[i];
                k = n_done[i];
#pragma unroll
                for (int j = 0; j < MM_TAN_NP; j++) {
                    Sz[j] = si[j];
                    Sv[j] = si[MM_TAN_NP + j];
                }
            }
            if (ch == 0 || (k == k0 && !(fabs(z) > z_escape)))
                tangent_steps_ge<GE, NG, NE>(P, z, v, k, k1, Sz, Sv, Bgrad, mass, inv_mass, dt, z_escape);
            zf[i] = z;
            vf[i] = v;
            n_done[i] = k;
#pragma unroll
            for (int j = 0; j < MM_TAN_NP; j++) {
                si[j] = Sz[j];
                si[MM_TAN_NP + j] = Sv[j];
            }
        }
        mm_queue_publish(Q, g, ch);
    }
}
