// Checkpointed reverse-mode gradient through the 1D integrator (the adjoint
// of rate_adjoint.hip's k_vjp_forward / k_vjp_segment): its argument block
// and, for the plan-specialised kernels of rate_spec.hip, the two kernel
// bodies over an evaluator EV of the force (mm_traj.h's interface) and GE of
// the force gradient (mm_tangent.h's).  The generic kernels in
// rate_adjoint.hip do the same steps with the tables in device memory.
// Free of host headers: run-time compiled kernels include it.
#pragma once

#include "mm_tangent.h"

struct VjpArgs {
    long n;
    int n_steps, K, n_seg;
    double Bgrad, mass, dt, z_escape;
    const double *z0, *v0, *cot_z, *cot_v;
    const int* n_up;
    double *zf, *vf, *gth, *lam, *ckpt, *jbuf;
    int* n_done;
};

// Forward pass: the state every K updates into the checkpoints, the final
// state, the initial adjoint (the cotangent) and zeroed parameter sums.
template <class EV>
__device__ __forceinline__ void vjp_forward_body(const MMRatePlanDev* __restrict__ P, const VjpArgs& A, long i) {
    if (i >= A.n) return;
    const double inv_mass = 1.0 / A.mass;
    double z = A.z0[i], v = A.v0[i];
    int lim = A.n_steps - 1;
    if (A.n_up && A.n_up[i] >= 0 && A.n_up[i] < lim) lim = A.n_up[i];
    double* ck = A.ckpt + i * 2L * (A.n_seg + 1);
    ck[0] = z;
    ck[1] = v;
    int k = 0;
    while (k < lim) {
        double f, r;
        EV::eval(P, v, z, A.Bgrad, f, r, nullptr, 64);
        const double a = mm_div_rn(f, A.mass, inv_mass);
        v = v + a * A.dt;
        z = z + v * A.dt;
        k++;
        if (k % A.K == 0) {
            ck[2 * (k / A.K)] = z;
            ck[2 * (k / A.K) + 1] = v;
        }
        if (fabs(z) > A.z_escape) break;
    }
    A.zf[i] = z;
    A.vf[i] = v;
    A.n_done[i] = k;
    A.lam[2 * i] = A.cot_z ? A.cot_z[i] : 0.0;
    A.lam[2 * i + 1] = A.cot_v ? A.cot_v[i] : 0.0;
#pragma unroll
    for (int j = 0; j < 7; j++) A.gth[i * 7 + j] = 0.0;
}

// Segment seg: recompute its updates from the checkpoint, storing the force
// Jacobian of each, then walk them backwards.  Adjoint of one update, with
// mu = lambda_v + dt lambda_z and c = dt / m:
//   lambda_z <- lambda_z + mu c dF/dz,  lambda_v <- mu (1 + c dF/dv),
//   g_theta += mu c dF/dtheta.
template <class GE, int NG, int NE>
__device__ __forceinline__ void vjp_segment_body(const MMRatePlanDev* __restrict__ P, const VjpArgs& A, int seg,
                                                 long i) {
    if (i >= A.n) return;
    const double inv_mass = 1.0 / A.mass;
    const int nd = A.n_done[i];
    const int k0 = seg * A.K;
    if (k0 >= nd) return;
    const int k1 = min(k0 + A.K, nd);
    const double* ck = A.ckpt + i * 2L * (A.n_seg + 1);
    double z = ck[2 * seg], v = ck[2 * seg + 1];
    double* jb = A.jbuf + i * 9L * A.K;
    for (int k = k0; k < k1; k++) {
        double f, J[MM_ADJ_NJ];
        if (!GE::grad(v, z, A.Bgrad, f, J)) {
            double fd, r;
            rate_eval_dense<NG, NE>(P, v, z, A.Bgrad, &fd, &r, nullptr);
            f = fd;
#pragma unroll
            for (int j = 0; j < MM_ADJ_NJ; j++) J[j] = __builtin_nan("");
        }
#pragma unroll
        for (int j = 0; j < MM_ADJ_NJ; j++) jb[(long)(k - k0) * 9 + j] = J[j];
        const double a = mm_div_rn(f, A.mass, inv_mass);
        v = v + a * A.dt;
        z = z + v * A.dt;
    }
    double lz = A.lam[2 * i], lv = A.lam[2 * i + 1];
    double g[7];
#pragma unroll
    for (int j = 0; j < 7; j++) g[j] = A.gth[i * 7 + j];
    const double c = A.dt / A.mass;
    for (int k = k1 - 1; k >= k0; k--) {
        const double* J = jb + (long)(k - k0) * 9;
        const double mu = lv + A.dt * lz;
        const double mc = mu * c;
#pragma unroll
        for (int j = 0; j < 7; j++) g[j] += mc * J[j];
        lz = lz + mc * J[MM_ADJ_J_Z];
        lv = mu + mc * J[MM_ADJ_J_V];
    }
    A.lam[2 * i] = lz;
    A.lam[2 * i + 1] = lv;
#pragma unroll
    for (int j = 0; j < 7; j++) A.gth[i * 7 + j] = g[j];
}
