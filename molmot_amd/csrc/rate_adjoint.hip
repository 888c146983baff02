// gfx950 kernel and C ABI for the per-point force gradient (rate_adjoint.h).
#include <hip/hip_runtime.h>

#include <math.h>

#include "../../include/molmot_hip.h"
#include "mm_errors.h"
#include "mm_plan_host.h"
#include "mm_queue.h"
#include "mm_spec.h"
#include "mm_vjp.h"
#include "rate_adjoint.h"

// Slab values per ground state for the trajectory-gradient kernels, which
// run two waves per SIMD (registers): three (1/(1 + S_g) kept for the reverse
// pass) while 8 blocks of 64 threads still fit a CU's 160 KiB of LDS, i.e.
// 3 * 12 * 64 doubles; two for larger systems.
template <int NG, bool CO>
constexpr int mm_grad_slab() {
    return (CO && 3 * NG * 64 * 8 * 8 <= 160 * 1024) ? 3 : 2;
}

template <int NG, int NE, bool UNIT, bool CO, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(3))) void k_rate_grad(const MMRatePlanDev* __restrict__ P, const double* __restrict__ v,
                                                  const double* __restrict__ z, long n, double Bgrad,
                                                  double* __restrict__ F, double* __restrict__ Jout) {
    __shared__ double lds[mm_adj_slab<NG, NE>() * BS];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double f, J[MM_ADJ_NJ];
    const bool ok = rate_grad_fast<NG, NE, UNIT, 1, 1, 1, 1, CO>(P, v[i], z[i], Bgrad, f, J, lds + threadIdx.x, BS);
    if (!ok) {
        // a dark ground state or a reducible excited chain: the force is
        // still defined (dense path) but this gradient is not provided
        double fd, r;  // temporaries: &f would keep f in scratch memory
        rate_eval_dense<NG, NE>(P, v[i], z[i], Bgrad, &fd, &r, nullptr);
        f = fd;
#pragma unroll
        for (int j = 0; j < MM_ADJ_NJ; j++) J[j] = __builtin_nan("");
    }
    if (F) F[i] = f;
#pragma unroll
    for (int j = 0; j < MM_ADJ_NJ; j++) Jout[i * MM_ADJ_NJ + j] = J[j];
}

template <int NG>
constexpr int grad_block() {
    return 64;
}

template <int NG, int NE>
static int launch_grad(const mm_rate_plan* pl, long n, const double* v, const double* z, double B, double* F,
                       double* J, hipStream_t s) {
    constexpr int block = grad_block<NG>();
    if (pl->host.unit_dirs && pl->host.comp_order)
        hipLaunchKernelGGL((k_rate_grad<NG, NE, true, true, block>), dim3(grid_for(n, block)), dim3(block), 0, s,
                           pl->dev, v, z, n, B, F, J);
    else if (pl->host.unit_dirs)
        hipLaunchKernelGGL((k_rate_grad<NG, NE, true, false, block>), dim3(grid_for(n, block)), dim3(block), 0, s,
                           pl->dev, v, z, n, B, F, J);
    else
        hipLaunchKernelGGL((k_rate_grad<NG, NE, false, false, block>), dim3(grid_for(n, block)), dim3(block), 0, s,
                           pl->dev, v, z, n, B, F, J);
    return mm_check_launch("k_rate_grad");
}

// Sizes with a gradient kernel.
#define MM_GRAD_SIZE_LIST(X) \
    X(12, 4)                 \
    X(36, 4)                 \
    X(2, 1)                  \
    X(3, 1)                  \
    X(4, 2)

extern "C" int mm_rate_eq_grad_points(const mm_rate_plan* pl, long n, const double* v, const double* z, double B,
                                      double* F, double* J, void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "rate_eq_grad_points: null plan");
    if (n < 0) return mm_fail(MM_EINVAL, "rate_eq_grad_points: negative size");
    if (n == 0) return MM_OK;
    if (!v || !z || !J) return mm_fail(MM_EINVAL, "rate_eq_grad_points: null pointer");
    if (pl->host.n_comp > MM_ADJ_NC)
        return mm_fail(MM_EUNSUPPORTED, "rate_eq_grad_points: at most %d frequency components (plan has %d)",
                       MM_ADJ_NC, pl->host.n_comp);
    if (pl->host.force_dense)
        return mm_fail(MM_EUNSUPPORTED, "rate_eq_grad_points: an excited state has no decay channel");
    if (mm_spec_grad_ready(pl)) return mm_spec_launch_grad_points(pl, v, z, n, B, F, J, (hipStream_t)stream);
    const int ng = pl->host.n_g, ne = pl->host.n_e;
#define MM_GR(G, E) \
    if (ng == G && ne == E) return launch_grad<G, E>(pl, n, v, z, B, F, J, (hipStream_t)stream);
    MM_GRAD_SIZE_LIST(MM_GR)
#undef MM_GR
    return mm_fail(MM_EUNSUPPORTED, "rate_eq_grad_points: no gradient kernel for (%d, %d)", ng, ne);
}

// ---------------------------------------------------------------------------
// Reverse-mode gradient through the 1D trajectory integrator (no recoil
// kicks): v_{k+1} = v_k + (F(v_k, z_k; theta) / m) dt, z_{k+1} = z_k + v_{k+1} dt,
// as in k_traj_euler (molmot/propagation/trajectories.py:19-75).  For the
// loss L = cot_z z_N + cot_v v_N at the last state N (the escape step, a
// requested stop step, or n_steps - 1) it returns dL/dtheta (7 entries per
// particle), dL/dz_0 and dL/dv_0.  Adjoint of one step, with
// mu = lambda_v + dt lambda_z and c = dt / m:
//   lambda_z <- lambda_z + mu c dF/dz,  lambda_v <- mu (1 + c dF/dv),
//   g_theta += mu c dF/dtheta.
// Checkpoints (z, v) every K states are written by the forward pass; each
// segment is then recomputed from its checkpoint with the force Jacobian
// stored per step in the workspace and walked backwards, last segment
// first.  Workspace per particle: 2 (ceil(N/K) + 1) doubles of checkpoints
// and 9 K doubles of Jacobians.
// ---------------------------------------------------------------------------

template <int NG, int NE, bool UNIT, int BS, bool PK>
__global__ __launch_bounds__(BS) void k_vjp_forward(const MMRatePlanDev* __restrict__ P, VjpArgs A) {
    __shared__ double lds[2 * NG * BS];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
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
        rate_eval<NG, NE, UNIT, 1, PK>(P, v, z, A.Bgrad, f, r, nullptr, lds + threadIdx.x, BS);
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

template <int NG, int NE, bool UNIT, bool CO, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2))) void k_vjp_segment(const MMRatePlanDev* __restrict__ P, VjpArgs A, int seg) {
    constexpr int SL = mm_grad_slab<NG, CO>();
    __shared__ double lds[mm_adj_slab<NG, NE, SL>() * BS];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n) return;
    const double inv_mass = 1.0 / A.mass;
    const int nd = A.n_done[i];
    const int k0 = seg * A.K;
    if (k0 >= nd) return;
    const int k1 = min(k0 + A.K, nd);
    const double* ck = A.ckpt + i * 2L * (A.n_seg + 1);
    double z = ck[2 * seg], v = ck[2 * seg + 1];
    double* jb = A.jbuf + i * 9L * A.K;
    // recompute the segment, storing the force Jacobian of every step
    for (int k = k0; k < k1; k++) {
        double f, J[MM_ADJ_NJ];
        if (!rate_grad_fast<NG, NE, UNIT, 4, 4, 4, 4, CO, SL>(P, v, z, A.Bgrad, f, J, lds + threadIdx.x, BS)) {
            double fd, r;  // temporaries: &f would keep f in scratch memory
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
    // walk the segment backwards
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

extern "C" long mm_traj_vjp_workspace_doubles(long n, int n_steps, int K) {
    if (n < 0 || n_steps < 1 || K < 1) return -1;
    const long n_seg = (n_steps - 1 + K - 1) / K;
    return n * (2 * (n_seg + 1) + 9L * K + 2);
}

template <int NG, int NE, bool UNIT, bool CO>
static int launch_vjp(const mm_rate_plan* pl, VjpArgs A, hipStream_t s) {
    const unsigned grid = (unsigned)((A.n + 63) / 64);
    // component-ordered records imply packed ones: the forward takes the
    // packed-plan variant with them; plan-specialised kernels where the plan
    // carries them (the same force bits; Jacobians equal to rounding)
    int rc;
    if (mm_spec_ready(pl)) {
        rc = mm_spec_launch_vjp_forward(pl, A, s);
    } else {
        hipLaunchKernelGGL((k_vjp_forward<NG, NE, UNIT, 64, CO>), dim3(grid), dim3(64), 0, s, pl->dev, A);
        rc = mm_check_launch("k_vjp_forward");
    }
    if (rc) return rc;
    const bool spec = mm_spec_grad_ready(pl);
    for (int seg = A.n_seg - 1; seg >= 0; seg--) {
        if (spec) {
            rc = mm_spec_launch_vjp_segment(pl, A, seg, s);
        } else {
            hipLaunchKernelGGL((k_vjp_segment<NG, NE, UNIT, CO, 64>), dim3(grid), dim3(64), 0, s, pl->dev, A, seg);
            rc = mm_check_launch("k_vjp_segment");
        }
        if (rc) return rc;
    }
    return MM_OK;
}

extern "C" int mm_traj_vjp_1d(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B,
                              double mass, double dt, int n_steps, const int* n_up, double z_escape, int K,
                              const double* cot_z, const double* cot_v, double* z_final, double* v_final,
                              int* n_done, double* g_theta, double* g_z0, double* g_v0, double* work,
                              long work_doubles, void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "traj_vjp_1d: null plan");
    if (n < 0 || n_steps < 1 || K < 1) return mm_fail(MM_EINVAL, "traj_vjp_1d: n >= 0, n_steps >= 1, K >= 1");
    if (!(mass > 0.0) || !(dt > 0.0)) return mm_fail(MM_EINVAL, "traj_vjp_1d: mass and dt must be positive");
    if (n == 0) return MM_OK;
    if (!z0 || !v0 || !z_final || !v_final || !n_done || !g_theta || !g_z0 || !g_v0 || !work)
        return mm_fail(MM_EINVAL, "traj_vjp_1d: null pointer");
    const long need = mm_traj_vjp_workspace_doubles(n, n_steps, K);
    if (work_doubles < need) return mm_fail(MM_EINVAL, "traj_vjp_1d: workspace %ld doubles, needs %ld", work_doubles, need);
    if (pl->host.n_comp > MM_ADJ_NC)
        return mm_fail(MM_EUNSUPPORTED, "traj_vjp_1d: at most %d frequency components", MM_ADJ_NC);
    if (pl->host.force_dense) return mm_fail(MM_EUNSUPPORTED, "traj_vjp_1d: an excited state has no decay channel");
    VjpArgs A;
    A.n = n;
    A.n_steps = n_steps;
    A.K = K;
    A.n_seg = (n_steps - 1 + K - 1) / K;
    A.Bgrad = B;
    A.mass = mass;
    A.dt = dt;
    A.z_escape = z_escape;
    A.z0 = z0;
    A.v0 = v0;
    A.cot_z = cot_z;
    A.cot_v = cot_v;
    A.n_up = n_up;
    A.zf = z_final;
    A.vf = v_final;
    A.gth = g_theta;
    A.n_done = n_done;
    A.ckpt = work;
    A.jbuf = work + n * 2L * (A.n_seg + 1);
    A.lam = A.jbuf + n * 9L * K;
    const int ng = pl->host.n_g, ne = pl->host.n_e;
    const bool unit = pl->host.unit_dirs != 0;
    int rc = MM_EUNSUPPORTED;
    bool found = false;
#define MM_VJ(G, E)                                                                         \
    if (ng == G && ne == E) {                                                               \
        found = true;                                                                       \
        rc = (unit && pl->host.comp_order) ? launch_vjp<G, E, true, true>(pl, A, (hipStream_t)stream) \
             : unit ? launch_vjp<G, E, true, false>(pl, A, (hipStream_t)stream)                      \
                    : launch_vjp<G, E, false, false>(pl, A, (hipStream_t)stream);                    \
    }
    MM_GRAD_SIZE_LIST(MM_VJ)
#undef MM_VJ
    if (!found) return mm_fail(MM_EUNSUPPORTED, "traj_vjp_1d: no gradient kernel for (%d, %d)", ng, ne);
    if (rc) return rc;
    // dL/dz0, dL/dv0 are the final adjoint state
    if (hipMemcpy2DAsync(g_z0, sizeof(double), A.lam, 2 * sizeof(double), sizeof(double), n,
                         hipMemcpyDeviceToDevice, (hipStream_t)stream) != hipSuccess ||
        hipMemcpy2DAsync(g_v0, sizeof(double), A.lam + 1, 2 * sizeof(double), sizeof(double), n,
                         hipMemcpyDeviceToDevice, (hipStream_t)stream) != hipSuccess)
        return mm_hip_fail("traj_vjp_1d: copy of the initial-state adjoint");
    return MM_OK;
}

// ---------------------------------------------------------------------------
// Forward sensitivities of the same integrator: one pass carries
// S_z = d z_k / d p and S_v = d v_k / d p for the nine parameters
// p = (det0..3, split, s0 scale, B', z_0, v_0), with the force Jacobian of
// each step from rate_grad_fast:
//   S_v <- S_v + c (dF/dz S_z + dF/dv S_v + dF/dp),   S_z <- S_z + dt S_v,
// c = dt / m (dF/dp is zero for z_0, v_0).  The result is the full 2 x 9
// Jacobian of the final state, so any cotangent's VJP is a contraction;
// nothing is stored per step and no segment is recomputed, so the cost is
// one gradient evaluation per step.  The escape step is treated as fixed,
// as in the checkpointed adjoint.
// ---------------------------------------------------------------------------
#define MM_TAN_NP (MM_ADJ_NJ)
// The tangent kernel's gradient loops over ground states are unrolled by 4
// (100k particles x 2,000 steps: 43.7 ms, against 44.1 ms unrolled by 2 and
// 45.1 ms by 6).
#ifndef MM_TAN_GU
#define MM_TAN_GU 4
#endif
#ifndef MM_TAN_GU1
#define MM_TAN_GU1 MM_TAN_GU
#endif
#ifndef MM_TAN_GU2
#define MM_TAN_GU2 MM_TAN_GU
#endif
#ifndef MM_TAN_GU3
#define MM_TAN_GU3 MM_TAN_GU
#endif
#ifndef MM_TAN_GU4
#define MM_TAN_GU4 MM_TAN_GU
#endif

// Updates k < k_end of one particle's state and 2 x 9 sensitivities; stops
// after the update that leaves |z| > z_escape.  The update of k_traj_tangent
// below, written out there again so that its register allocation stays as
// measured; the queue test checks the two bit for bit.
template <int NG, int NE, bool UNIT, bool CO, int BS, int SL>
__device__ __forceinline__ void tangent_steps(const MMRatePlanDev* __restrict__ P, double& z, double& v, int& k,
                                              int k_end, double (&Sz)[MM_TAN_NP], double (&Sv)[MM_TAN_NP],
                                              double Bgrad, double mass, double inv_mass, double dt,
                                              double z_escape, double* slab) {
    const double c = dt / mass;
    while (k < k_end) {
        double f, J[MM_ADJ_NJ];
        if (!rate_grad_fast<NG, NE, UNIT, MM_TAN_GU1, MM_TAN_GU2, MM_TAN_GU3, MM_TAN_GU4, CO, SL>(P, v, z, Bgrad, f,
                                                                                                 J, slab, BS)) {
            double fd, r;  // temporaries: &f would keep f in scratch memory
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

template <int NG, int NE, bool UNIT, bool CO, int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2))) void k_traj_tangent(
    const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,
    double Bgrad, double mass, double dt, int n_steps, const int* __restrict__ n_up, double z_escape,
    double* __restrict__ zf, double* __restrict__ vf, int* __restrict__ n_done, double* __restrict__ sens) {
    constexpr int SL = mm_grad_slab<NG, CO>();
    __shared__ double lds[mm_adj_slab<NG, NE, SL>() * BS];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double inv_mass = 1.0 / mass;
    double z = z0[i], v = v0[i];
    int lim = n_steps - 1;
    if (n_up && n_up[i] >= 0 && n_up[i] < lim) lim = n_up[i];
    // columns 0..6: theta, 7: z_0, 8: v_0
    double Sz[MM_TAN_NP], Sv[MM_TAN_NP];
#pragma unroll
    for (int j = 0; j < MM_TAN_NP; j++) {
        Sz[j] = (j == MM_ADJ_NJ - 2) ? 1.0 : 0.0;
        Sv[j] = (j == MM_ADJ_NJ - 1) ? 1.0 : 0.0;
    }
    const double c = dt / mass;
    int k = 0;
    while (k < lim) {
        double f, J[MM_ADJ_NJ];
        if (!rate_grad_fast<NG, NE, UNIT, MM_TAN_GU1, MM_TAN_GU2, MM_TAN_GU3, MM_TAN_GU4, CO, SL>(P, v, z, Bgrad, f,
                                                                                                 J, lds + threadIdx.x,
                                                                                                 BS)) {
            double fd, r;  // temporaries: &f would keep f in scratch memory
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
    zf[i] = z;
    vf[i] = v;
    n_done[i] = k;
#pragma unroll
    for (int j = 0; j < MM_TAN_NP; j++) {
        sens[i * 2 * MM_TAN_NP + j] = Sz[j];
        sens[i * 2 * MM_TAN_NP + MM_TAN_NP + j] = Sv[j];
    }
}

// Persistent form of k_traj_tangent on the queue of mm_queue.h: a particle's
// (z, v, updates done, sensitivities) live in the output arrays between
// chunks of updates; it continues while it has run every earlier update
// and is still inside z_escape.  Same arithmetic per update, so the results
// are bit-identical to k_traj_tangent's.
template <int NG, int NE, bool UNIT, bool CO>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_traj_tangent_queue(
    const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,
    double Bgrad, double mass, double dt, int n_steps, const int* __restrict__ n_up, double z_escape,
    double* __restrict__ zf, double* __restrict__ vf, int* __restrict__ n_done, double* __restrict__ sens,
    MMQueue Q) {
    constexpr int SL = mm_grad_slab<NG, CO>();
    __shared__ double lds[mm_adj_slab<NG, NE, SL>() * 64];
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
                tangent_steps<NG, NE, UNIT, CO, 64, SL>(P, z, v, k, k1, Sz, Sv, Bgrad, mass, inv_mass, dt, z_escape,
                                                        lds + threadIdx.x);
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

extern "C" int mm_traj_tangent_1d(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B,
                                  double mass, double dt, int n_steps, const int* n_up, double z_escape,
                                  double* z_final, double* v_final, int* n_done, double* sens, void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "traj_tangent_1d: null plan");
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "traj_tangent_1d: n >= 0, n_steps >= 1");
    if (!(mass > 0.0) || !(dt > 0.0)) return mm_fail(MM_EINVAL, "traj_tangent_1d: mass and dt must be positive");
    if (n == 0) return MM_OK;
    if (!z0 || !v0 || !z_final || !v_final || !n_done || !sens)
        return mm_fail(MM_EINVAL, "traj_tangent_1d: null pointer");
    if (pl->host.n_comp > MM_ADJ_NC)
        return mm_fail(MM_EUNSUPPORTED, "traj_tangent_1d: at most %d frequency components", MM_ADJ_NC);
    if (pl->host.force_dense) return mm_fail(MM_EUNSUPPORTED, "traj_tangent_1d: an excited state has no decay channel");
    const int ng = pl->host.n_g, ne = pl->host.n_e;
    const bool unit = pl->host.unit_dirs != 0;
    const unsigned grid = (unsigned)((n + 63) / 64);
    hipStream_t s = (hipStream_t)stream;
    if (mm_spec_grad_ready(pl))
        return mm_spec_launch_tangent_grid(pl, n, z0, v0, B, mass, dt, n_steps, n_up, z_escape, z_final, v_final,
                                           n_done, sens, s);
#define MM_TG(G, E)                                                                                          \
    if (ng == G && ne == E) {                                                                                \
        if (unit && pl->host.comp_order)                                                                     \
            hipLaunchKernelGGL((k_traj_tangent<G, E, true, true, 64>), dim3(grid), dim3(64), 0, s, pl->dev, n, z0, \
                               v0, B, mass, dt, n_steps, n_up, z_escape, z_final, v_final, n_done, sens);    \
        else if (unit)                                                                                       \
            hipLaunchKernelGGL((k_traj_tangent<G, E, true, false, 64>), dim3(grid), dim3(64), 0, s, pl->dev, n,  \
                               z0, v0, B, mass, dt, n_steps, n_up, z_escape, z_final, v_final, n_done, sens); \
        else                                                                                                 \
            hipLaunchKernelGGL((k_traj_tangent<G, E, false, false, 64>), dim3(grid), dim3(64), 0, s, pl->dev, n, \
                               z0, v0, B, mass, dt, n_steps, n_up, z_escape, z_final, v_final, n_done, sens); \
        return mm_check_launch("k_traj_tangent");                                                            \
    }
    MM_GRAD_SIZE_LIST(MM_TG)
#undef MM_TG
    return mm_fail(MM_EUNSUPPORTED, "traj_tangent_1d: no gradient kernel for (%d, %d)", ng, ne);
}

template <int NG, int NE, bool UNIT, bool CO>
static int launch_tangent_queue(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B,
                                double mass, double dt, int n_steps, const int* n_up, double z_escape,
                                double* z_final, double* v_final, int* n_done, double* sens, int schedule, void* ws,
                                long ws_bytes, hipStream_t s) {
    MMQueueShape q;
    const bool can = ws != nullptr && mm_queue_shape(n, (long)n_steps - 1, q) && ws_bytes >= q.bytes;
    if (schedule == 2 && !can)
        return mm_fail(MM_EINVAL, "traj_tangent_1d_ws: the queue needs n_steps > %d and a workspace of "
                                  "mm_traj_euler_workspace_bytes() bytes", 2 * MM_TQ_MIN_CHUNK);
    const bool spec = mm_spec_grad_ready(pl);
    if (schedule != 1 && can) {
        MMQueueSlots sl;
        const int rc = spec ? mm_spec_tangent_slots(pl, sl)
                            : mm_queue_slots((const void*)k_traj_tangent_queue<NG, NE, UNIT, CO>, sl);
        if (rc != MM_OK) return rc;
        if (schedule == 2 || mm_queue_pays(q, sl)) {
            unsigned grid;
            size_t dyn;
            mm_queue_grid(q, sl, &grid, &dyn);
            const MMQueue Q = mm_queue_make(ws, q, sl);
            const long words = q.bytes / 8;
            hipLaunchKernelGGL(k_mm_queue_clear, dim3(grid_for(words, 256)), dim3(256), 0, s,
                               (unsigned long long*)ws, words);
            if (spec) {
                const int lrc = mm_spec_launch_tangent_queue(pl, grid, dyn, n, z0, v0, B, mass, dt, n_steps, n_up,
                                                             z_escape, z_final, v_final, n_done, sens, Q, s);
                if (lrc != MM_OK) return lrc;
            } else {
                hipLaunchKernelGGL((k_traj_tangent_queue<NG, NE, UNIT, CO>), dim3(grid), dim3(64), dyn, s, pl->dev,
                                   n, z0, v0, B, mass, dt, n_steps, n_up, z_escape, z_final, v_final, n_done, sens,
                                   Q);
            }
            hipLaunchKernelGGL(k_mm_queue_status, dim3(grid_for(n, 256)), dim3(256), 0, s, (const unsigned*)Q.ctl,
                               n_done, n);
            return mm_check_launch("k_traj_tangent_queue");
        }
    }
    if (spec)
        return mm_spec_launch_tangent_grid(pl, n, z0, v0, B, mass, dt, n_steps, n_up, z_escape, z_final, v_final,
                                           n_done, sens, s);
    hipLaunchKernelGGL((k_traj_tangent<NG, NE, UNIT, CO, 64>), dim3(grid_for(n, 64)), dim3(64), 0, s, pl->dev, n,
                       z0, v0, B, mass, dt, n_steps, n_up, z_escape, z_final, v_final, n_done, sens);
    return mm_check_launch("k_traj_tangent");
}

extern "C" int mm_traj_tangent_1d_ws(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B,
                                     double mass, double dt, int n_steps, const int* n_up, double z_escape,
                                     double* z_final, double* v_final, int* n_done, double* sens, int schedule,
                                     void* workspace, long workspace_bytes, void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "traj_tangent_1d_ws: null plan");
    if (schedule < 0 || schedule > 2)
        return mm_fail(MM_EINVAL, "traj_tangent_1d_ws: schedule is 0 (auto), 1 (grid) or 2 (queue)");
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "traj_tangent_1d_ws: n >= 0, n_steps >= 1");
    if (!(mass > 0.0) || !(dt > 0.0)) return mm_fail(MM_EINVAL, "traj_tangent_1d_ws: mass and dt must be positive");
    if (n == 0) return MM_OK;
    if (!z0 || !v0 || !z_final || !v_final || !n_done || !sens)
        return mm_fail(MM_EINVAL, "traj_tangent_1d_ws: null pointer");
    if (workspace_bytes < 0) return mm_fail(MM_EINVAL, "traj_tangent_1d_ws: workspace_bytes >= 0 required");
    if (workspace && ((uintptr_t)workspace & 255u))
        return mm_fail(MM_EINVAL, "traj_tangent_1d_ws: the workspace must be 256-byte aligned");
    if (pl->host.n_comp > MM_ADJ_NC)
        return mm_fail(MM_EUNSUPPORTED, "traj_tangent_1d_ws: at most %d frequency components", MM_ADJ_NC);
    if (pl->host.force_dense)
        return mm_fail(MM_EUNSUPPORTED, "traj_tangent_1d_ws: an excited state has no decay channel");
    const int ng = pl->host.n_g, ne = pl->host.n_e;
    const bool unit = pl->host.unit_dirs != 0;
    hipStream_t s = (hipStream_t)stream;
#define MM_TGQ(G, E)                                                                                          \
    if (ng == G && ne == E) {                                                                                 \
        if (unit && pl->host.comp_order)                                                                      \
            return launch_tangent_queue<G, E, true, true>(pl, n, z0, v0, B, mass, dt, n_steps, n_up, z_escape, \
                                                          z_final, v_final, n_done, sens, schedule, workspace,  \
                                                          workspace_bytes, s);                                 \
        if (unit)                                                                                             \
            return launch_tangent_queue<G, E, true, false>(pl, n, z0, v0, B, mass, dt, n_steps, n_up, z_escape, \
                                                           z_final, v_final, n_done, sens, schedule, workspace,  \
                                                           workspace_bytes, s);                                 \
        return launch_tangent_queue<G, E, false, false>(pl, n, z0, v0, B, mass, dt, n_steps, n_up, z_escape,    \
                                                        z_final, v_final, n_done, sens, schedule, workspace,     \
                                                        workspace_bytes, s);                                    \
    }
    MM_GRAD_SIZE_LIST(MM_TGQ)
#undef MM_TGQ
    return mm_fail(MM_EUNSUPPORTED, "traj_tangent_1d_ws: no gradient kernel for (%d, %d)", ng, ne);
}
