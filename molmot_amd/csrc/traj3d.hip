// gfx950 kernels and C ABI: 3D per-axis trajectories and Monte-Carlo
// temperature statistics on the 1D rate-equation force (rate_core.h).
#include <hip/hip_runtime.h>

#include <math.h>

#include "../../include/molmot_hip.h"
#include "mm_errors.h"
#include "mm_plan_host.h"
#include "mm_axis3d.h"
#include "mm_queue.h"
#include "mm_spec.h"
#include "mm_vsq.h"
#include "rate_core.h"

// Per-axis 3D model (molmot/propagation/trajectories.py:78-156; the lane-quad
// step of mm_axis3d.h over the tables in device memory).
template <int NG, int NE, bool UNIT, bool SAVE, int BS, bool RF = false, bool PK = false>
__global__ __launch_bounds__(BS) void k_traj_axis3d(const MMRatePlanDev* __restrict__ P, long n,
                                                    const double* __restrict__ r0, const double* __restrict__ v0,
                                                    double Bgrad, double mass, double dt, int n_steps,
                                                    double r_escape, double* __restrict__ rt,
                                                    double* __restrict__ vt, double* __restrict__ rf,
                                                    double* __restrict__ vf, int* __restrict__ nact,
                                                    const MMRatePlanDev* __restrict__ P2 = nullptr) {
    __shared__ double lds[2 * NG * BS];
    axis3d_grid_body<MMEvalPlan<NG, NE, UNIT, 1, PK>, NG, NE, UNIT, SAVE, BS, RF>(
        P, P2, (long)blockIdx.x * blockDim.x + threadIdx.x, n, r0, v0, Bgrad, mass, dt, n_steps, r_escape, rt, vt, rf,
        vf, nact, lds + threadIdx.x);
}

// Persistent form of k_traj_axis3d (no saved trajectory, no RF) on the queue
// of mm_queue.h; the queue test checks it against k_traj_axis3d bit for bit.
template <int NG, int NE, bool UNIT, bool PK>
__global__ __launch_bounds__(64) void k_traj_axis3d_queue(const MMRatePlanDev* __restrict__ P, long n,
                                                          const double* __restrict__ r0,
                                                          const double* __restrict__ v0, double Bgrad, double mass,
                                                          double dt, int n_steps, double r_escape,
                                                          double* __restrict__ rf, double* __restrict__ vf,
                                                          int* __restrict__ nact, MMQueue Q) {
    __shared__ double lds[2 * NG * 64];
    axis3d_queue_body<MMEvalPlan<NG, NE, UNIT, 1, PK>>(P, n, r0, v0, Bgrad, mass, dt, n_steps, r_escape, rf, vf, nact,
                                                       Q, lds + threadIdx.x);
}

// Monte-Carlo temperature statistics (molmot/propagation/trajectories.py:159-229;
// the step loop is mm_vsq.h's, over the tables in device memory).
template <int NG, int NE, bool UNIT, int BS, bool RF = false, bool PK = false>
__global__ __launch_bounds__(BS) void k_traj_vsq(const MMRatePlanDev* __restrict__ P, long n,
                                                 const double* __restrict__ z0, const double* __restrict__ v0,
                                                 double Bgrad, double mass, double dt, int n_steps, int i_eq,
                                                 double z_escape, double* __restrict__ vsq,
                                                 long long* __restrict__ cnt,
                                                 const MMRatePlanDev* __restrict__ P2 = nullptr) {
    __shared__ double lds[2 * NG * BS];
    vsq_grid_body<MMEvalPlan<NG, NE, UNIT, 1, PK>, NG, NE, UNIT, BS, RF>(
        P, P2, (long)blockIdx.x * blockDim.x + threadIdx.x, n, z0, v0, Bgrad, mass, dt, n_steps, i_eq, z_escape, vsq,
        cnt, lds + threadIdx.x);
}

// Persistent form of k_traj_vsq (no RF) on the queue of mm_queue.h; the
// queue test checks it against k_traj_vsq bit for bit.
template <int NG, int NE, bool UNIT, bool PK>
__global__ __launch_bounds__(64) void k_traj_vsq_queue(const MMRatePlanDev* __restrict__ P, long n,
                                                       const double* __restrict__ z0, const double* __restrict__ v0,
                                                       double Bgrad, double mass, double dt, int n_steps, int i_eq,
                                                       double z_escape, double* __restrict__ vsq,
                                                       long long* __restrict__ cnt, double* __restrict__ zs,
                                                       double* __restrict__ vs, MMQueue Q) {
    __shared__ double lds[2 * NG * 64];
    vsq_queue_body<MMEvalPlan<NG, NE, UNIT, 1, PK>, NG, NE, UNIT>(P, n, z0, v0, Bgrad, mass, dt, n_steps, i_eq,
                                                                  z_escape, vsq, cnt, zs, vs, Q, lds + threadIdx.x);
}

// after k_traj_vsq_queue: a timed-out queue wait leaves every count -1
static __global__ void k_vsq_queue_status(const unsigned* __restrict__ ctl, long long* __restrict__ cnt, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && ctl[2] != 0u) cnt[i] = -1;
}

template <int NG, int NE>
static int launch_axis3d(const mm_rate_plan* pl, long n, const double* r0, const double* v0, double B, double m,
                         double dt, int ns, double re, double* rt, double* vt, double* rf, double* vf, int* na,
                         hipStream_t s, const mm_rate_plan* p2 = nullptr) {
    constexpr int block = 64;
    const long threads = 4 * n;
    const bool save = rt != nullptr;
    const bool unit = pl->host.unit_dirs != 0;
    if (p2) {  // RF: beams along +-z (checked by the caller)
        if (save)
            hipLaunchKernelGGL((k_traj_axis3d<NG, NE, true, true, block, true>), dim3(grid_for(threads, block)),
                               dim3(block), 0, s, pl->dev, n, r0, v0, B, m, dt, ns, re, rt, vt, rf, vf, na, p2->dev);
        else
            hipLaunchKernelGGL((k_traj_axis3d<NG, NE, true, false, block, true>), dim3(grid_for(threads, block)),
                               dim3(block), 0, s, pl->dev, n, r0, v0, B, m, dt, ns, re, rt, vt, rf, vf, na, p2->dev);
        return mm_check_launch("k_traj_axis3d (RF)");
    }
    if (!save && mm_spec_ready(pl)) return mm_spec_launch_axis3d_grid(pl, n, r0, v0, B, m, dt, ns, re, rf, vf, na, s);
    // packed plans take the variant without the general pair path
    const bool packed = unit && pl->host.packed;
#define MM_A3(U, S, PK)                                                                                          \
    if (unit == U && save == S && packed == PK)                                                                  \
        hipLaunchKernelGGL((k_traj_axis3d<NG, NE, U, S, block, false, PK>), dim3(grid_for(threads, block)),       \
                           dim3(block), 0, s, pl->dev, n, r0, v0, B, m, dt, ns, re, rt, vt, rf, vf, na);
    MM_A3(true, false, true)
    MM_A3(true, true, true)
    MM_A3(true, false, false)
    MM_A3(true, true, false)
    MM_A3(false, false, false)
    MM_A3(false, true, false)
#undef MM_A3
    return mm_check_launch("k_traj_axis3d");
}

template <int NG, int NE>
static int launch_vsq(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                      double dt, int ns, int ieq, double ze, double* vsq, long long* cnt, hipStream_t s,
                      const mm_rate_plan* p2 = nullptr) {
    if (!p2 && mm_spec_ready(pl)) return mm_spec_launch_vsq_grid(pl, n, z0, v0, B, m, dt, ns, ieq, ze, vsq, cnt, s);
    constexpr int block = 64;
    if (p2 && mm_spec_rf_ready(pl, p2)) return mm_spec_launch_rf_vsq(pl, p2, n, z0, v0, B, m, dt, ns, ieq, ze, vsq, cnt, s);
    if (p2) {  // RF: beams along +-z (checked by the caller)
        hipLaunchKernelGGL((k_traj_vsq<NG, NE, true, block, true>), dim3(grid_for(n, block)), dim3(block), 0, s,
                           pl->dev, n, z0, v0, B, m, dt, ns, ieq, ze, vsq, cnt, p2->dev);
        return mm_check_launch("k_traj_vsq (RF)");
    }
    if (pl->host.unit_dirs && pl->host.packed)
        hipLaunchKernelGGL((k_traj_vsq<NG, NE, true, block, false, true>), dim3(grid_for(n, block)), dim3(block), 0,
                           s, pl->dev, n, z0, v0, B, m, dt, ns, ieq, ze, vsq, cnt);
    else if (pl->host.unit_dirs)
        hipLaunchKernelGGL((k_traj_vsq<NG, NE, true, block>), dim3(grid_for(n, block)), dim3(block), 0, s, pl->dev,
                           n, z0, v0, B, m, dt, ns, ieq, ze, vsq, cnt);
    else
        hipLaunchKernelGGL((k_traj_vsq<NG, NE, false, block>), dim3(grid_for(n, block)), dim3(block), 0, s,
                           pl->dev, n, z0, v0, B, m, dt, ns, ieq, ze, vsq, cnt);
    return mm_check_launch("k_traj_vsq");
}

#define MM_TRAJ_SIZE_LIST(X) \
    X(12, 4)                 \
    X(2, 1)                  \
    X(3, 1)

template <int NG, int NE>
static int launch_axis3d_sched(const mm_rate_plan* pl, long n, const double* r0, const double* v0, double B, double m,
                               double dt, int ns, double re, double* rt, double* vt, double* rf, double* vf, int* na,
                               int schedule, void* ws, long ws_bytes, hipStream_t s) {
    MMQueueShape q;  // 4 lanes per particle: groups of 16 particles
    const bool can = rt == nullptr && ws != nullptr && mm_queue_shape(4 * n, (long)ns - 1, q) && ws_bytes >= q.bytes;
    if (schedule == 2 && !can)
        return mm_fail(MM_EINVAL, "traj_axis3d_ws: the queue needs no saved trajectory, n_steps > %d and a "
                                  "workspace of mm_traj_axis3d_workspace_bytes() bytes", 2 * MM_TQ_MIN_CHUNK);
    if (schedule != 1 && can) {
        const bool unit = pl->host.unit_dirs != 0;
        const bool packed = unit && pl->host.packed;
        const bool spec = mm_spec_ready(pl);
        const void* kern = packed ? (const void*)k_traj_axis3d_queue<NG, NE, true, true>
                           : unit ? (const void*)k_traj_axis3d_queue<NG, NE, true, false>
                                  : (const void*)k_traj_axis3d_queue<NG, NE, false, false>;
        MMQueueSlots sl;
        const int rc = spec ? mm_spec_axis3d_slots(pl, sl) : mm_queue_slots(kern, sl);
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
                const int lrc = mm_spec_launch_axis3d_queue(pl, grid, dyn, n, r0, v0, B, m, dt, ns, re, rf, vf, na, Q,
                                                            s);
                if (lrc != MM_OK) return lrc;
            } else if (packed)
                hipLaunchKernelGGL((k_traj_axis3d_queue<NG, NE, true, true>), dim3(grid), dim3(64), dyn, s, pl->dev,
                                   n, r0, v0, B, m, dt, ns, re, rf, vf, na, Q);
            else if (unit)
                hipLaunchKernelGGL((k_traj_axis3d_queue<NG, NE, true, false>), dim3(grid), dim3(64), dyn, s, pl->dev,
                                   n, r0, v0, B, m, dt, ns, re, rf, vf, na, Q);
            else
                hipLaunchKernelGGL((k_traj_axis3d_queue<NG, NE, false, false>), dim3(grid), dim3(64), dyn, s,
                                   pl->dev, n, r0, v0, B, m, dt, ns, re, rf, vf, na, Q);
            hipLaunchKernelGGL(k_mm_queue_status, dim3(grid_for(n, 256)), dim3(256), 0, s, (const unsigned*)Q.ctl,
                               na, n);
            return mm_check_launch("k_traj_axis3d_queue");
        }
    }
    return launch_axis3d<NG, NE>(pl, n, r0, v0, B, m, dt, ns, re, rt, vt, rf, vf, na, s);
}

extern "C" long mm_traj_axis3d_workspace_bytes(long n, int n_steps) {
    MMQueueShape q;
    return mm_queue_shape(4 * n, (long)n_steps - 1, q) ? q.bytes : 0;
}

extern "C" int mm_traj_axis3d_ws(const mm_rate_plan* pl, long n, const double* r0, const double* v0, double B,
                                 double mass, double dt, int n_steps, double r_escape, double* r_traj, double* v_traj,
                                 double* r_final, double* v_final, int* n_actual, int schedule, void* workspace,
                                 long workspace_bytes, void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "traj_axis3d_ws: null plan");
    if (schedule < 0 || schedule > 2)
        return mm_fail(MM_EINVAL, "traj_axis3d_ws: schedule is 0 (auto), 1 (grid) or 2 (queue)");
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "traj_axis3d_ws: n >= 0 and n_steps >= 1 required");
    if (n == 0) return MM_OK;
    if (!r0 || !v0 || !r_final || !v_final || !n_actual) return mm_fail(MM_EINVAL, "traj_axis3d_ws: null pointer");
    if ((r_traj == nullptr) != (v_traj == nullptr))
        return mm_fail(MM_EINVAL, "traj_axis3d_ws: r_traj and v_traj go together");
    if (!(mass > 0.0)) return mm_fail(MM_EINVAL, "traj_axis3d_ws: mass must be positive");
    if (workspace_bytes < 0) return mm_fail(MM_EINVAL, "traj_axis3d_ws: workspace_bytes >= 0 required");
    if (workspace && ((uintptr_t)workspace & 255u))
        return mm_fail(MM_EINVAL, "traj_axis3d_ws: the workspace must be 256-byte aligned");
    const int ng = pl->host.n_g, ne = pl->host.n_e;
#define MM_T3Q(G, E)                                                                                            \
    if (ng == G && ne == E)                                                                                     \
        return launch_axis3d_sched<G, E>(pl, n, r0, v0, B, mass, dt, n_steps, r_escape, r_traj, v_traj, r_final, \
                                         v_final, n_actual, schedule, workspace, workspace_bytes,               \
                                         (hipStream_t)stream);
    MM_TRAJ_SIZE_LIST(MM_T3Q)
#undef MM_T3Q
    return mm_fail(MM_EUNSUPPORTED, "traj_axis3d_ws: no kernel for (%d, %d)", ng, ne);
}

extern "C" int mm_traj_axis3d(const mm_rate_plan* pl, long n, const double* r0, const double* v0, double B,
                              double mass, double dt, int n_steps, double r_escape, double* r_traj, double* v_traj,
                              double* r_final, double* v_final, int* n_actual, void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "traj_axis3d: null plan");
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "traj_axis3d: n >= 0 and n_steps >= 1 required");
    if (n == 0) return MM_OK;
    if (!r0 || !v0 || !r_final || !v_final || !n_actual) return mm_fail(MM_EINVAL, "traj_axis3d: null pointer");
    if ((r_traj == nullptr) != (v_traj == nullptr)) return mm_fail(MM_EINVAL, "traj_axis3d: r_traj and v_traj go together");
    if (!(mass > 0.0)) return mm_fail(MM_EINVAL, "traj_axis3d: mass must be positive");
    const int ng = pl->host.n_g, ne = pl->host.n_e;
#define MM_T3(G, E)                                                                                            \
    if (ng == G && ne == E)                                                                                    \
        return launch_axis3d<G, E>(pl, n, r0, v0, B, mass, dt, n_steps, r_escape, r_traj, v_traj, r_final,    \
                                   v_final, n_actual, (hipStream_t)stream);
    MM_TRAJ_SIZE_LIST(MM_T3)
#undef MM_T3
    return mm_fail(MM_EUNSUPPORTED, "traj_axis3d: no kernel for (%d, %d)", ng, ne);
}

// workspace of the queued temperature run: the queue, then z and v of
// every particle
static long vsq_queue_bytes(const MMQueueShape& q, long n) { return ((q.bytes + 255) & ~255l) + 16 * n; }

template <int NG, int NE>
static int launch_vsq_sched(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                            double dt, int ns, int ieq, double ze, double* vsq, long long* cnt, int schedule,
                            void* ws, long ws_bytes, hipStream_t s) {
    MMQueueShape q;
    const bool can = ws != nullptr && mm_queue_shape(n, (long)ns - 1, q) && ws_bytes >= vsq_queue_bytes(q, n);
    if (schedule == 2 && !can)
        return mm_fail(MM_EINVAL, "traj_vsq_1d_ws: the queue needs n_steps > %d and a workspace of "
                                  "mm_traj_vsq_workspace_bytes() bytes", 2 * MM_TQ_MIN_CHUNK);
    if (schedule != 1 && can) {
        const bool unit = pl->host.unit_dirs != 0;
        const bool packed = unit && pl->host.packed;
        const bool spec = mm_spec_ready(pl);
        const void* kern = packed ? (const void*)k_traj_vsq_queue<NG, NE, true, true>
                           : unit ? (const void*)k_traj_vsq_queue<NG, NE, true, false>
                                  : (const void*)k_traj_vsq_queue<NG, NE, false, false>;
        MMQueueSlots sl;
        const int rc = spec ? mm_spec_vsq_slots(pl, sl) : mm_queue_slots(kern, sl);
        if (rc != MM_OK) return rc;
        if (schedule == 2 || mm_queue_pays(q, sl)) {
            unsigned grid;
            size_t dyn;
            mm_queue_grid(q, sl, &grid, &dyn);
            const MMQueue Q = mm_queue_make(ws, q, sl);
            double* zs = (double*)((char*)ws + ((q.bytes + 255) & ~255l));
            double* vs = zs + n;
            const long words = q.bytes / 8;
            hipLaunchKernelGGL(k_mm_queue_clear, dim3(grid_for(words, 256)), dim3(256), 0, s,
                               (unsigned long long*)ws, words);
            if (spec) {
                const int lrc = mm_spec_launch_vsq_queue(pl, grid, dyn, n, z0, v0, B, m, dt, ns, ieq, ze, vsq, cnt, zs,
                                                         vs, Q, s);
                if (lrc != MM_OK) return lrc;
            } else if (packed)
                hipLaunchKernelGGL((k_traj_vsq_queue<NG, NE, true, true>), dim3(grid), dim3(64), dyn, s, pl->dev, n,
                                   z0, v0, B, m, dt, ns, ieq, ze, vsq, cnt, zs, vs, Q);
            else if (unit)
                hipLaunchKernelGGL((k_traj_vsq_queue<NG, NE, true, false>), dim3(grid), dim3(64), dyn, s, pl->dev, n,
                                   z0, v0, B, m, dt, ns, ieq, ze, vsq, cnt, zs, vs, Q);
            else
                hipLaunchKernelGGL((k_traj_vsq_queue<NG, NE, false, false>), dim3(grid), dim3(64), dyn, s, pl->dev,
                                   n, z0, v0, B, m, dt, ns, ieq, ze, vsq, cnt, zs, vs, Q);
            // a timed-out wait: the sample counts become -1
            hipLaunchKernelGGL(k_vsq_queue_status, dim3(grid_for(n, 256)), dim3(256), 0, s, (const unsigned*)Q.ctl,
                               cnt, n);
            return mm_check_launch("k_traj_vsq_queue");
        }
    }
    return launch_vsq<NG, NE>(pl, n, z0, v0, B, m, dt, ns, ieq, ze, vsq, cnt, s);
}

extern "C" long mm_traj_vsq_workspace_bytes(long n, int n_steps) {
    MMQueueShape q;
    return mm_queue_shape(n, (long)n_steps - 1, q) ? vsq_queue_bytes(q, n) : 0;
}

extern "C" int mm_traj_vsq_1d_ws(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B,
                                 double mass, double dt, int n_steps, int i_eq, double z_escape, double* vsq_sum,
                                 long long* n_samples, int schedule, void* workspace, long workspace_bytes,
                                 void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "traj_vsq_1d_ws: null plan");
    if (schedule < 0 || schedule > 2)
        return mm_fail(MM_EINVAL, "traj_vsq_1d_ws: schedule is 0 (auto), 1 (grid) or 2 (queue)");
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "traj_vsq_1d_ws: n >= 0 and n_steps >= 1 required");
    if (n == 0) return MM_OK;
    if (!z0 || !v0 || !vsq_sum || !n_samples) return mm_fail(MM_EINVAL, "traj_vsq_1d_ws: null pointer");
    if (!(mass > 0.0)) return mm_fail(MM_EINVAL, "traj_vsq_1d_ws: mass must be positive");
    if (workspace_bytes < 0) return mm_fail(MM_EINVAL, "traj_vsq_1d_ws: workspace_bytes >= 0 required");
    if (workspace && ((uintptr_t)workspace & 255u))
        return mm_fail(MM_EINVAL, "traj_vsq_1d_ws: the workspace must be 256-byte aligned");
    const int ng = pl->host.n_g, ne = pl->host.n_e;
#define MM_TVQ(G, E)                                                                                            \
    if (ng == G && ne == E)                                                                                     \
        return launch_vsq_sched<G, E>(pl, n, z0, v0, B, mass, dt, n_steps, i_eq, z_escape, vsq_sum, n_samples,  \
                                      schedule, workspace, workspace_bytes, (hipStream_t)stream);
    MM_TRAJ_SIZE_LIST(MM_TVQ)
#undef MM_TVQ
    return mm_fail(MM_EUNSUPPORTED, "traj_vsq_1d_ws: no kernel for (%d, %d)", ng, ne);
}

extern "C" int mm_traj_vsq_1d(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B,
                              double mass, double dt, int n_steps, int i_eq, double z_escape, double* vsq_sum,
                              long long* n_samples, void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "traj_vsq_1d: null plan");
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "traj_vsq_1d: n >= 0 and n_steps >= 1 required");
    if (n == 0) return MM_OK;
    if (!z0 || !v0 || !vsq_sum || !n_samples) return mm_fail(MM_EINVAL, "traj_vsq_1d: null pointer");
    if (!(mass > 0.0)) return mm_fail(MM_EINVAL, "traj_vsq_1d: mass must be positive");
    const int ng = pl->host.n_g, ne = pl->host.n_e;
#define MM_TV(G, E)                                                                                         \
    if (ng == G && ne == E)                                                                                 \
        return launch_vsq<G, E>(pl, n, z0, v0, B, mass, dt, n_steps, i_eq, z_escape, vsq_sum, n_samples,    \
                                (hipStream_t)stream);
    MM_TRAJ_SIZE_LIST(MM_TV)
#undef MM_TV
    return mm_fail(MM_EUNSUPPORTED, "traj_vsq_1d: no kernel for (%d, %d)", ng, ne);
}

static int rf_pair_check(const mm_rate_plan* p0, const mm_rate_plan* p1, const char* who) {
    if (!p0 || !p1) return mm_fail(MM_EINVAL, "%s: null plan", who);
    if (p0->host.n_g != p1->host.n_g || p0->host.n_e != p1->host.n_e)
        return mm_fail(MM_EINVAL, "%s: the two phase plans differ in size", who);
    if (!p0->host.unit_dirs || !p1->host.unit_dirs) return mm_fail(MM_EUNSUPPORTED, "%s: beams must be along +-z", who);
    return MM_OK;
}

extern "C" int mm_traj_axis3d_rf(const mm_rate_plan* p0, const mm_rate_plan* p1, long n, const double* r0,
                                 const double* v0, double B, double mass, double dt, int n_steps, double r_escape,
                                 double* r_traj, double* v_traj, double* r_final, double* v_final, int* n_actual,
                                 void* stream) {
    mm_clear_error();
    int rc = rf_pair_check(p0, p1, "traj_axis3d_rf");
    if (rc) return rc;
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "traj_axis3d_rf: n >= 0 and n_steps >= 1 required");
    if (n == 0) return MM_OK;
    if (!r0 || !v0 || !r_final || !v_final || !n_actual) return mm_fail(MM_EINVAL, "traj_axis3d_rf: null pointer");
    if ((r_traj == nullptr) != (v_traj == nullptr))
        return mm_fail(MM_EINVAL, "traj_axis3d_rf: r_traj and v_traj go together");
    if (!(mass > 0.0)) return mm_fail(MM_EINVAL, "traj_axis3d_rf: mass must be positive");
    const int ng = p0->host.n_g, ne = p0->host.n_e;
#define MM_T3RF(G, E)                                                                                         \
    if (ng == G && ne == E)                                                                                   \
        return launch_axis3d<G, E>(p0, n, r0, v0, B, mass, dt, n_steps, r_escape, r_traj, v_traj, r_final,   \
                                   v_final, n_actual, (hipStream_t)stream, p1);
    MM_TRAJ_SIZE_LIST(MM_T3RF)
#undef MM_T3RF
    return mm_fail(MM_EUNSUPPORTED, "traj_axis3d_rf: no kernel for (%d, %d)", ng, ne);
}

extern "C" int mm_traj_vsq_rf_1d(const mm_rate_plan* p0, const mm_rate_plan* p1, long n, const double* z0,
                                 const double* v0, double B, double mass, double dt, int n_steps, int i_eq,
                                 double z_escape, double* vsq_sum, long long* n_samples, void* stream) {
    mm_clear_error();
    int rc = rf_pair_check(p0, p1, "traj_vsq_rf_1d");
    if (rc) return rc;
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "traj_vsq_rf_1d: n >= 0 and n_steps >= 1 required");
    if (n == 0) return MM_OK;
    if (!z0 || !v0 || !vsq_sum || !n_samples) return mm_fail(MM_EINVAL, "traj_vsq_rf_1d: null pointer");
    if (!(mass > 0.0)) return mm_fail(MM_EINVAL, "traj_vsq_rf_1d: mass must be positive");
    const int ng = p0->host.n_g, ne = p0->host.n_e;
#define MM_TVRF(G, E)                                                                                       \
    if (ng == G && ne == E)                                                                                 \
        return launch_vsq<G, E>(p0, n, z0, v0, B, mass, dt, n_steps, i_eq, z_escape, vsq_sum, n_samples,    \
                                (hipStream_t)stream, p1);
    MM_TRAJ_SIZE_LIST(MM_TVRF)
#undef MM_TVRF
    return mm_fail(MM_EUNSUPPORTED, "traj_vsq_rf_1d: no kernel for (%d, %d)", ng, ne);
}
