// gfx950 kernels and C ABI for the 1D rate-equation force and trajectories.
// See include/molmot_hip.h for the contract of each entry point.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/molmot_hip.h"
#include "mm_errors.h"
#include "mm_plan.h"
#include "mm_queue.h"
#include "mm_traj.h"
#include "philox.h"
#include "rate_core.h"

#include "mm_plan_host.h"
#include "mm_spec.h"

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

template <int NG, int NE, bool UNIT, int BS, bool PK = false>
__global__ __launch_bounds__(BS) void k_rate_map(const MMRatePlanDev* __restrict__ P,
                                                            const double* __restrict__ v,
                                                            const double* __restrict__ z, int nz, long n,
                                                            double Bgrad, double* __restrict__ F,
                                                            double* __restrict__ pop, double* __restrict__ R) {
    __shared__ double lds[2 * NG * BS];
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    const long iv = idx / nz;
    const long iz = idx - iv * nz;
    double f, r;
    rate_eval<NG, NE, UNIT, 1, PK>(P, v[iv], z[iz], Bgrad, f, r, pop ? pop + idx * NG : nullptr, lds + threadIdx.x,
                                   BS);
    F[idx] = f;
    if (R) R[idx] = r;
}

template <int NG, int NE, bool UNIT, int BS, bool PK = false>
__global__ __launch_bounds__(BS) void k_rate_points(const MMRatePlanDev* __restrict__ P,
                                                               const double* __restrict__ v,
                                                               const double* __restrict__ z, long n, double Bgrad,
                                                               double* __restrict__ F, double* __restrict__ pop,
                                                               double* __restrict__ R) {
    __shared__ double lds[2 * NG * BS];
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    double f, r;
    rate_eval<NG, NE, UNIT, 1, PK>(P, v[idx], z[idx], Bgrad, f, r, pop ? pop + idx * NG : nullptr, lds + threadIdx.x,
                                   BS);
    F[idx] = f;
    if (R) R[idx] = r;
}

// One thread per particle (mm_traj.h).
template <int NG, int NE, bool UNIT, bool SAVE, bool KICKS, int BS, bool RF = false, bool PK = false>
__global__ __launch_bounds__(BS) void k_traj_euler(const MMRatePlanDev* __restrict__ P, long n,
                                                    const double* __restrict__ z0, const double* __restrict__ v0,
                                                    double Bgrad, double mass, double dt, int n_steps,
                                                    double z_escape, double* __restrict__ zt,
                                                    double* __restrict__ vt, double* __restrict__ zf,
                                                    double* __restrict__ vf, int* __restrict__ nact,
                                                    double kick_scale, uint32_t k0, uint32_t k1, uint32_t sid,
                                                    const MMRatePlanDev* __restrict__ P2 = nullptr) {
    __shared__ double lds[2 * NG * BS];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    traj_grid_body<MMEvalPlan<NG, NE, UNIT, MM_TRAJ_GU, PK>, NG, NE, UNIT, SAVE, KICKS, BS, RF>(
        P, i, n, z0, v0, Bgrad, mass, dt, n_steps, z_escape, zt, vt, zf, vf, nact, kick_scale, k0, k1, sid, P2,
        lds + threadIdx.x);
}

// Persistent work-queue form (mm_traj.h, mm_queue_dev.h).
template <int NG, int NE, bool UNIT, bool KICKS, bool PK>
__global__ __launch_bounds__(64) void k_traj_euler_queue(
    const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,
    double Bgrad, double mass, double dt, int n_steps, double z_escape, double* __restrict__ zf,
    double* __restrict__ vf, int* __restrict__ nact, double kick_scale, uint32_t k0, uint32_t k1, uint32_t sid,
    MMQueue Q) {
    __shared__ double lds[2 * NG * 64];
    traj_queue_body<MMEvalPlan<NG, NE, UNIT, MM_TRAJ_GU, PK>, NG, NE, UNIT, KICKS>(
        P, n, z0, v0, Bgrad, mass, dt, n_steps, z_escape, zf, vf, nact, kick_scale, k0, k1, sid, Q,
        lds + threadIdx.x);
}

// Ensemble statistics at sample steps (molmot/mot/analysis.py:128-288; the
// body, MMMoments and the merge are mm_traj.h's, here over the tables in
// device memory).
template <int NG, int NE, bool UNIT, bool KICKS, bool PK>
__global__ __launch_bounds__(64) void k_traj_stats(const MMRatePlanDev* __restrict__ P, long n,
                                                   const double* __restrict__ z0, const double* __restrict__ v0,
                                                   double Bgrad, double mass, double dt, int n_steps,
                                                   double z_escape, double r_alive, const int* __restrict__ samp,
                                                   int n_samp, double* __restrict__ part, double kick_scale,
                                                   uint32_t k0, uint32_t k1, uint32_t sid) {
    __shared__ double lds[2 * NG * 64];
    traj_stats_body<MMEvalPlan<NG, NE, UNIT, MM_TRAJ_GU, PK>, NG, NE, UNIT, KICKS>(
        P, n, z0, v0, Bgrad, mass, dt, n_steps, z_escape, r_alive, samp, n_samp, part, kick_scale, k0, k1, sid,
        lds + threadIdx.x);
}

// one thread per sample: the block partials merged in block order; out[k] =
// (alive count, mean z, sum (z - mean z)^2, mean v, sum (v - mean v)^2)
__global__ void k_traj_stats_reduce(const double* __restrict__ part, long n_blocks, int n_samp,
                                    double* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_samp) return;
    MMMoments a = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (long b = 0; b < n_blocks; b++) {
        const double* q = part + (b * n_samp + k) * 5;
        const MMMoments m = {q[0], q[1], q[2], q[3], q[4]};
        mm_moments_merge(a, m);
    }
    out[k * 5] = a.n;
    out[k * 5 + 1] = a.mz;
    out[k * 5 + 2] = a.m2z;
    out[k * 5 + 3] = a.mv;
    out[k * 5 + 4] = a.m2v;
}

template <int NG, int NE>
static int launch_traj_stats(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B,
                             double m, double dt, int ns, double ze, double ra, const int* samp, int n_samp,
                             double* part, double* out, double ks, uint32_t k0, uint32_t k1, uint32_t sid,
                             hipStream_t s) {
    const bool unit = pl->host.unit_dirs != 0, packed = unit && pl->host.packed, kicks = ks > 0.0;
    const dim3 grid((unsigned)((n + 63) / 64)), block(64);
    if (mm_spec_ready(pl)) {
        const int rc = mm_spec_launch_stats(pl, n, z0, v0, B, m, dt, ns, ze, ra, samp, n_samp, part, ks, k0, k1, sid, s);
        if (rc != MM_OK) return rc;
        hipLaunchKernelGGL(k_traj_stats_reduce, dim3((n_samp + 63) / 64), dim3(64), 0, s, (const double*)part,
                           (long)grid.x, n_samp, out);
        return mm_check_launch("k_traj_stats_reduce");
    }
#define MM_TS(U, K, PKV)                                                                                      \
    hipLaunchKernelGGL((k_traj_stats<NG, NE, U, K, PKV>), grid, block, 0, s, pl->dev, n, z0, v0, B, m, dt, ns, \
                       ze, ra, samp, n_samp, part, ks, k0, k1, sid)
    if (packed) {
        if (kicks) MM_TS(true, true, true); else MM_TS(true, false, true);
    } else if (unit) {
        if (kicks) MM_TS(true, true, false); else MM_TS(true, false, false);
    } else {
        if (kicks) MM_TS(false, true, false); else MM_TS(false, false, false);
    }
#undef MM_TS
    hipLaunchKernelGGL(k_traj_stats_reduce, dim3((n_samp + 63) / 64), dim3(64), 0, s, (const double*)part,
                       (long)grid.x, n_samp, out);
    return mm_check_launch("k_traj_stats");
}

__global__ void k_philox(long n, const uint32_t* __restrict__ ctr, uint32_t k0, uint32_t k1,
                         uint32_t* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const mm_u32x4 c = {ctr[4 * i], ctr[4 * i + 1], ctr[4 * i + 2], ctr[4 * i + 3]};
    const mm_u32x4 o = mm_philox4x32_10(c, k0, k1);
    out[4 * i] = o.x;
    out[4 * i + 1] = o.y;
    out[4 * i + 2] = o.z;
    out[4 * i + 3] = o.w;
}

// ---------------------------------------------------------------------------
// size dispatch: (n_ground, n_excited) pairs instantiated for this build
// ---------------------------------------------------------------------------


// Block size of the map/points kernels: the plan's setting, else 256
// threads, fewer when the per-thread LDS slab (2*NG doubles) would exceed
// 48 KiB per block.
static bool block_fits(int ng, int block) { return 2 * ng * 8 * block <= 49152; }

static int map_block(const mm_rate_plan* pl) {
    if (pl->block) return pl->block;
    const int ng = pl->host.n_g;
    return block_fits(ng, 256) ? 256 : block_fits(ng, 128) ? 128 : 64;
}

template <int NG, int NE, int BS>
static void launch_map_bs(const mm_rate_plan* pl, const double* v, const double* z, int nz, long n, double B,
                          double* F, double* pop, double* R, hipStream_t s) {
    if constexpr (2 * NG * 8 * BS <= 49152) {
        if (pl->host.unit_dirs && pl->host.packed)
            hipLaunchKernelGGL((k_rate_map<NG, NE, true, BS, true>), dim3(grid_for(n, BS)), dim3(BS), 0, s, pl->dev, v,
                               z, nz, n, B, F, pop, R);
        else if (pl->host.unit_dirs)
            hipLaunchKernelGGL((k_rate_map<NG, NE, true, BS>), dim3(grid_for(n, BS)), dim3(BS), 0, s, pl->dev, v, z,
                               nz, n, B, F, pop, R);
        else
            hipLaunchKernelGGL((k_rate_map<NG, NE, false, BS>), dim3(grid_for(n, BS)), dim3(BS), 0, s, pl->dev, v,
                               z, nz, n, B, F, pop, R);
    }
}

template <int NG, int NE, int BS>
static void launch_points_bs(const mm_rate_plan* pl, const double* v, const double* z, long n, double B, double* F,
                             double* pop, double* R, hipStream_t s) {
    if constexpr (2 * NG * 8 * BS <= 49152) {
        if (pl->host.unit_dirs && pl->host.packed)
            hipLaunchKernelGGL((k_rate_points<NG, NE, true, BS, true>), dim3(grid_for(n, BS)), dim3(BS), 0, s, pl->dev,
                               v, z, n, B, F, pop, R);
        else if (pl->host.unit_dirs)
            hipLaunchKernelGGL((k_rate_points<NG, NE, true, BS>), dim3(grid_for(n, BS)), dim3(BS), 0, s, pl->dev, v,
                               z, n, B, F, pop, R);
        else
            hipLaunchKernelGGL((k_rate_points<NG, NE, false, BS>), dim3(grid_for(n, BS)), dim3(BS), 0, s, pl->dev,
                               v, z, n, B, F, pop, R);
    }
}

template <int NG, int NE>
static int launch_map(const mm_rate_plan* pl, const double* v, const double* z, int nz, long n, double B,
                      double* F, double* pop, double* R, hipStream_t s) {
    const int block = map_block(pl);
    if (block == 64) launch_map_bs<NG, NE, 64>(pl, v, z, nz, n, B, F, pop, R, s);
    else if (block == 128) launch_map_bs<NG, NE, 128>(pl, v, z, nz, n, B, F, pop, R, s);
    else launch_map_bs<NG, NE, 256>(pl, v, z, nz, n, B, F, pop, R, s);
    return mm_check_launch("k_rate_map");
}

template <int NG, int NE>
static int launch_points(const mm_rate_plan* pl, const double* v, const double* z, long n, double B, double* F,
                         double* pop, double* R, hipStream_t s) {
    const int block = map_block(pl);
    if (block == 64) launch_points_bs<NG, NE, 64>(pl, v, z, n, B, F, pop, R, s);
    else if (block == 128) launch_points_bs<NG, NE, 128>(pl, v, z, n, B, F, pop, R, s);
    else launch_points_bs<NG, NE, 256>(pl, v, z, n, B, F, pop, R, s);
    return mm_check_launch("k_rate_points");
}

extern "C" int mm_rate_plan_set_block(mm_rate_plan* pl, int block) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "rate_plan_set_block: null plan");
    if (block != 0 && block != 64 && block != 128 && block != 256)
        return mm_fail(MM_EINVAL, "rate_plan_set_block: block is 0 (default), 64, 128 or 256");
    if (block && !block_fits(pl->host.n_g, block))
        return mm_fail(MM_EINVAL, "rate_plan_set_block: %d threads x %d-double LDS slab exceed 48 KiB", block,
                       2 * pl->host.n_g);
    pl->block = block;
    return MM_OK;
}

extern "C" int mm_rate_plan_get_block(const mm_rate_plan* pl) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "rate_plan_get_block: null plan");
    return map_block(pl);
}

template <int NG, int NE, bool UNIT, bool SAVE, bool KICKS, bool PK>
static void launch_traj_t(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                          double dt, int ns, double ze, double* zt, double* vt, double* zf, double* vf, int* na,
                          double ks, uint32_t k0, uint32_t k1, uint32_t sid, hipStream_t s) {
    // one wavefront per block: an ensemble of 1e5 particles gives ~1600
    // blocks, spread evenly over the 256 CUs
    constexpr int block = 64;
    hipLaunchKernelGGL((k_traj_euler<NG, NE, UNIT, SAVE, KICKS, block, false, PK>), dim3(grid_for(n, block)),
                       dim3(block), 0, s, pl->dev, n, z0, v0, B, m, dt, ns, ze, zt, vt, zf, vf, na, ks, k0, k1, sid);
}

template <int NG, int NE>
static int launch_traj(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                       double dt, int ns, double ze, double* zt, double* vt, double* zf, double* vf, int* na,
                       double ks, uint32_t k0, uint32_t k1, uint32_t sid, hipStream_t s) {
    // on a plan with specialised kernels (rate_spec.hip): the same
    // trajectories bit for bit
    if (mm_spec_ready(pl))
        return zt == nullptr
                   ? mm_spec_launch_grid(pl, n, z0, v0, B, m, dt, ns, ze, zf, vf, na, ks, k0, k1, sid, s)
                   : mm_spec_launch_grid_save(pl, n, z0, v0, B, m, dt, ns, ze, zt, vt, zf, vf, na, ks, k0, k1, sid, s);
    const bool unit = pl->host.unit_dirs != 0;
    const bool packed = unit && pl->host.packed;
    const bool save = zt != nullptr;
    const bool kicks = ks > 0.0;
#define MM_TRAJ_CASE(U, S, K, PK)                                                                        \
    if (unit == U && packed == PK && save == S && kicks == K)                                            \
        launch_traj_t<NG, NE, U, S, K, PK>(pl, n, z0, v0, B, m, dt, ns, ze, zt, vt, zf, vf, na, ks, k0, k1, sid, s);
    MM_TRAJ_CASE(true, false, false, true)
    MM_TRAJ_CASE(true, false, true, true)
    MM_TRAJ_CASE(true, true, false, true)
    MM_TRAJ_CASE(true, true, true, true)
    MM_TRAJ_CASE(true, false, false, false)
    MM_TRAJ_CASE(true, false, true, false)
    MM_TRAJ_CASE(true, true, false, false)
    MM_TRAJ_CASE(true, true, true, false)
    MM_TRAJ_CASE(false, false, false, false)
    MM_TRAJ_CASE(false, false, true, false)
    MM_TRAJ_CASE(false, true, false, false)
    MM_TRAJ_CASE(false, true, true, false)
#undef MM_TRAJ_CASE
    return mm_check_launch("k_traj_euler");
}

// ---- queued launches (mm_queue.h) ----
// the queue kernel a plan and kick setting run, as a pointer for the
// occupancy query
template <int NG, int NE>
static const void* traj_queue_kernel(const mm_rate_plan* pl, bool kicks) {
    const bool unit = pl->host.unit_dirs != 0;
    const bool packed = unit && pl->host.packed;
    if (packed) return kicks ? (const void*)k_traj_euler_queue<NG, NE, true, true, true>
                             : (const void*)k_traj_euler_queue<NG, NE, true, false, true>;
    if (unit) return kicks ? (const void*)k_traj_euler_queue<NG, NE, true, true, false>
                           : (const void*)k_traj_euler_queue<NG, NE, true, false, false>;
    return kicks ? (const void*)k_traj_euler_queue<NG, NE, false, true, false>
                 : (const void*)k_traj_euler_queue<NG, NE, false, false, false>;
}

template <int NG, int NE>
static int launch_traj_sched(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                             double dt, int ns, double ze, double* zt, double* vt, double* zf, double* vf, int* na,
                             double ks, uint32_t k0, uint32_t k1, uint32_t sid, int schedule, void* ws, long ws_bytes,
                             hipStream_t s) {
    MMQueueShape q;
    const bool can = zt == nullptr && ws != nullptr && mm_queue_shape(n, (long)ns - 1, q) && ws_bytes >= q.bytes;
    if (schedule == 2 && !can)
        return mm_fail(MM_EINVAL, "traj_euler_1d_ws: the queue needs no saved trajectory, n_steps > %d and a "
                                  "workspace of mm_traj_euler_workspace_bytes() bytes", 2 * MM_TQ_MIN_CHUNK);
    if (schedule != 1 && can) {
        const bool kicks = ks > 0.0;
        const bool spec = mm_spec_ready(pl);
        MMQueueSlots sl;
        const int rc = spec ? mm_spec_queue_slots(pl, kicks, sl)
                            : mm_queue_slots(traj_queue_kernel<NG, NE>(pl, kicks), sl);
        if (rc != MM_OK) return rc;
        if (schedule == 2 || mm_queue_pays(q, sl)) {
            unsigned grid;
            size_t dyn;
            mm_queue_grid(q, sl, &grid, &dyn);
            const MMQueue Q = mm_queue_make(ws, q, sl);
            const long words = q.bytes / 8;
            hipLaunchKernelGGL(k_mm_queue_clear, dim3(grid_for(words, 256)), dim3(256), 0, s,
                               (unsigned long long*)ws, words);
            const bool unit = pl->host.unit_dirs != 0;
            const bool packed = unit && pl->host.packed;
            if (spec) {
                const int lrc = mm_spec_launch_queue(pl, grid, dyn, n, z0, v0, B, m, dt, ns, ze, zf, vf, na, ks, k0,
                                                     k1, sid, Q, s);
                if (lrc != MM_OK) return lrc;
            }
#define MM_TQ_CASE(U, K, PKD)                                                                               \
    if (!spec && unit == U && kicks == K && packed == PKD)                                                  \
        hipLaunchKernelGGL((k_traj_euler_queue<NG, NE, U, K, PKD>), dim3(grid), dim3(64), dyn, s, pl->dev, n, \
                           z0, v0, B, m, dt, ns, ze, zf, vf, na, ks, k0, k1, sid, Q);
            MM_TQ_CASE(true, true, true)
            MM_TQ_CASE(true, false, true)
            MM_TQ_CASE(true, true, false)
            MM_TQ_CASE(true, false, false)
            MM_TQ_CASE(false, true, false)
            MM_TQ_CASE(false, false, false)
#undef MM_TQ_CASE
            hipLaunchKernelGGL(k_mm_queue_status, dim3(grid_for(n, 256)), dim3(256), 0, s, (const unsigned*)Q.ctl,
                               na, n);
            return mm_check_launch("k_traj_euler_queue");
        }
    }
    return launch_traj<NG, NE>(pl, n, z0, v0, B, m, dt, ns, ze, zt, vt, zf, vf, na, ks, k0, k1, sid, s);
}

template <int NG, int NE>
static int traj_auto_schedule(const mm_rate_plan* pl, long n, int ns, bool save, double ks) {
    MMQueueShape q;
    if (save || !mm_queue_shape(n, (long)ns - 1, q)) return 1;
    MMQueueSlots sl;
    const int rc = mm_spec_ready(pl) ? mm_spec_queue_slots(pl, ks > 0.0, sl)
                                     : mm_queue_slots(traj_queue_kernel<NG, NE>(pl, ks > 0.0), sl);
    if (rc != MM_OK) return rc;
    return mm_queue_pays(q, sl) ? 2 : 1;
}

template <int NG, int NE, bool SAVE, bool KICKS>
static void launch_traj_rf_t(const mm_rate_plan* p0, const mm_rate_plan* p1, long n, const double* z0,
                             const double* v0, double B, double m, double dt, int ns, double ze, double* zt,
                             double* vt, double* zf, double* vf, int* na, double ks, uint32_t k0, uint32_t k1,
                             uint32_t sid, hipStream_t s) {
    constexpr int block = 64;
    hipLaunchKernelGGL((k_traj_euler<NG, NE, true, SAVE, KICKS, block, true>), dim3(grid_for(n, block)), dim3(block),
                       0, s, p0->dev, n, z0, v0, B, m, dt, ns, ze, zt, vt, zf, vf, na, ks, k0, k1, sid, p1->dev);
}

template <int NG, int NE>
static int launch_traj_rf(const mm_rate_plan* p0, const mm_rate_plan* p1, long n, const double* z0,
                          const double* v0, double B, double m, double dt, int ns, double ze, double* zt, double* vt,
                          double* zf, double* vf, int* na, double ks, uint32_t k0, uint32_t k1, uint32_t sid,
                          hipStream_t s) {
    if (mm_spec_rf_ready(p0, p1))
        return mm_spec_launch_rf_traj(p0, p1, n, z0, v0, B, m, dt, ns, ze, zt, vt, zf, vf, na, ks, k0, k1, sid, s);
    const bool save = zt != nullptr, kicks = ks > 0.0;
    if (save && kicks) launch_traj_rf_t<NG, NE, true, true>(p0, p1, n, z0, v0, B, m, dt, ns, ze, zt, vt, zf, vf, na, ks, k0, k1, sid, s);
    else if (save) launch_traj_rf_t<NG, NE, true, false>(p0, p1, n, z0, v0, B, m, dt, ns, ze, zt, vt, zf, vf, na, ks, k0, k1, sid, s);
    else if (kicks) launch_traj_rf_t<NG, NE, false, true>(p0, p1, n, z0, v0, B, m, dt, ns, ze, zt, vt, zf, vf, na, ks, k0, k1, sid, s);
    else launch_traj_rf_t<NG, NE, false, false>(p0, p1, n, z0, v0, B, m, dt, ns, ze, zt, vt, zf, vf, na, ks, k0, k1, sid, s);
    return mm_check_launch("k_traj_euler (RF)");
}

// ---------------------------------------------------------------------------
// host-side plan construction
// ---------------------------------------------------------------------------

static int build_plan_host(MMRatePlanDev& P, int n_g, int n_e, int n_b, const double* E, const double* zd,
                           const double* d2, const double* dirs, const double* freqs, const int* pols,
                           const double* s0s, const int* comp, const double* split_sign, int n_comp,
                           double Gamma, double k, double Geff, double hbar) {
    memset(&P, 0, sizeof(P));
    if (n_g < 1 || n_g > MM_MAX_G || n_e < 1 || n_e > MM_MAX_E || n_b < 0 || n_b > MM_MAX_B)
        return mm_fail(MM_EINVAL, "plan: n_ground in [1,%d], n_excited in [1,%d], n_beams in [0,%d] required "
                                  "(got %d, %d, %d)",
                       MM_MAX_G, MM_MAX_E, MM_MAX_B, n_g, n_e, n_b);
    if (n_comp < 0 || n_comp > MM_MAX_COMP) return mm_fail(MM_EINVAL, "plan: n_comp must be in [0,%d]", MM_MAX_COMP);
    if (!E || !zd || !d2 || (n_b > 0 && (!dirs || !freqs || !pols || !s0s)))
        return mm_fail(MM_EINVAL, "plan: null table pointer");
    if (!(Gamma > 0.0) || !(k > 0.0)) return mm_fail(MM_EINVAL, "plan: Gamma and k must be positive");
    const int n_s = n_g + n_e;
    P.n_g = n_g;
    P.n_e = n_e;
    P.n_b = n_b;
    P.n_comp = n_comp;
    P.Gamma = Gamma;
    P.k = k;
    P.Geff = Geff;
    P.hbar = hbar;
    P.gh = Gamma / 2.0;
    P.ghs = P.gh * P.gh;
    P.gq = P.ghs / (MM_TWO_PI * MM_TWO_PI);
    P.mk = -k;
    P.gh_geff = P.gh * Geff;
    P.fk_gh_geff = hbar * k * P.gh * Geff;
    P.inv_gh_geff = P.gh_geff != 0.0 ? 1.0 / P.gh_geff : 0.0;
    for (int i = 0; i < n_s; i++) {
        P.E[i] = E[i];
        P.zd[i] = zd[i];
    }
    for (int g = 0; g < n_g; g++)
        for (int e = 0; e < n_e; e++)
            for (int q = 0; q < 3; q++) P.d2[g][e][q] = d2[(g * n_e + e) * 3 + q];
    // branching ratios in the reference's summation order
    // (molmot/obe/rate_equations_jit.py:37-49)
    P.force_dense = 0;
    for (int e = 0; e < n_e; e++) {
        double total = 0.0;
        for (int g = 0; g < n_g; g++)
            for (int q = 0; q < 3; q++) total += P.d2[g][e][q];
        if (total > 1e-30) {
            for (int g = 0; g < n_g; g++) {
                double s = 0.0;
                for (int q = 0; q < 3; q++) s += P.d2[g][e][q];
                P.BR[e][g] = s / total;
            }
        } else {
            P.force_dense = 1;  // an excited state that never decays: not a closed chain
        }
    }
    P.unit_dirs = 1;
    for (int b = 0; b < n_b; b++) {
        if (pols[b] < 0 || pols[b] > 2) return mm_fail(MM_EINVAL, "plan: beam %d polarisation %d not in {0,1,2}", b, pols[b]);
        P.bfreq[b] = freqs[b];
        P.bkdir[b] = dirs[b];
        P.bs0[b] = s0s[b];
        P.bpol[b] = pols[b];
        P.bcomp[b] = comp ? comp[b] : 0;
        if (comp && (comp[b] < 0 || comp[b] >= (n_comp > 0 ? n_comp : MM_MAX_COMP)))
            return mm_fail(MM_EINVAL, "plan: beam %d component %d out of range", b, comp[b]);
        if (fabs(dirs[b]) != 1.0) P.unit_dirs = 0;
    }
    for (int c = 0; c < n_comp; c++) P.split_sign[c] = split_sign ? split_sign[c] : 0.0;
    // coupling terms sorted by (g, e) pair; within a pair, with unit
    // directions, the +z beams come first (input order kept within a sign),
    // and a pair of exactly two +z and two -z terms is flagged for the
    // four-term path of mm_pair_sums
    int nt = 0;
    for (int g = 0; g < n_g; g++) {
        for (int e = 0; e < n_e; e++) {
            const int pair = g * n_e + e;
            P.pair_start[pair] = nt;
            for (int pass = 0; pass < (P.unit_dirs ? 2 : 1); pass++) {
                for (int b = 0; b < n_b; b++) {
                    if (P.unit_dirs && ((pass == 0) != (dirs[b] > 0.0))) continue;
                    const double dd = P.d2[g][e][pols[b]];
                    if (dd < 1e-15) continue;  // reference skip rule
                    if (nt >= MM_MAX_TERMS)
                        return mm_fail(MM_EINVAL, "plan: more than %d coupling terms", MM_MAX_TERMS);
                    MMTerm& T = P.terms[nt++];
                    T.freq = freqs[b];
                    T.kdir = dirs[b];
                    T.C2 = s0s[b] * dd * P.gq;
                    T.beam = b;
                    T.comp = P.bcomp[b];
                }
            }
            const int c = nt - P.pair_start[pair];
            const MMTerm* T = P.terms + P.pair_start[pair];
            P.pair_kind[pair] = (P.unit_dirs && c == 4 && T[0].kdir > 0 && T[1].kdir > 0 && T[2].kdir < 0 &&
                                 T[3].kdir < 0) ? 1 : 0;
        }
    }
    P.pair_start[n_g * n_e] = nt;
    P.n_terms = nt;
    // packed records: every pair empty or (+z, +z, -z, -z), n_e <= 4
    P.packed = (P.unit_dirs && n_e <= 4) ? 1 : 0;
    for (int pr = 0; pr < n_g * n_e && P.packed; pr++)
        if (P.pair_start[pr + 1] != P.pair_start[pr] && !P.pair_kind[pr]) P.packed = 0;
    if (P.packed) {
        for (int g = 0; g < n_g; g++) {
            MMGroundRec& G = P.grec[g];
            G.E = P.E[g];
            G.zd = P.zd[g];
            for (int f = 0; f < 4; f++) G.br[f] = f < n_e ? P.BR[f][g] : 0.0;
            G.mask = G.dirm = 0;
            G.Gamma = P.Gamma;
            P.pair_mask[g] = 0;
            for (int e = 0; e < n_e; e++) {
                const int pr = g * n_e + e;
                MMPairRec& Q = P.prec[g * 4 + e];
                for (int j = 0; j < 4; j++) Q.f[j] = Q.c2[j] = 0.0;
                if (!P.pair_kind[pr]) continue;
                P.pair_mask[g] |= 1u << e;
                G.mask = P.pair_mask[g];
                for (int j = 0; j < 4; j++) {
                    Q.f[j] = P.terms[P.pair_start[pr] + j].freq;
                    Q.c2[j] = P.terms[P.pair_start[pr] + j].C2;
                }
            }
        }
        // component-ordered records: every packed pair must hold one term of
        // each component 0..3
        P.comp_order = 1;
        for (int g = 0; g < n_g && P.comp_order; g++)
            for (int e = 0; e < n_e; e++) {
                if (!(P.pair_mask[g] & (1u << e))) continue;
                unsigned seen = 0;
                for (int j = 0; j < 4; j++) {
                    const int c = P.terms[P.pair_start[g * n_e + e] + j].comp;
                    if (c >= 0 && c < 4) seen |= 1u << c;
                }
                if (seen != 0xfu) P.comp_order = 0;
            }
        for (int g = 0; g < n_g && P.comp_order; g++) {
            for (int e = 0; e < n_e; e++) {
                MMPairRec& Q = P.precc[g * 4 + e];
                for (int j = 0; j < 4; j++) Q.f[j] = Q.c2[j] = 0.0;
                if (!(P.pair_mask[g] & (1u << e))) continue;
                for (int j = 0; j < 4; j++) {
                    const MMTerm& T = P.terms[P.pair_start[g * n_e + e] + j];
                    Q.f[T.comp] = T.freq;
                    Q.c2[T.comp] = T.C2;
                    if (T.kdir < 0) P.grec[g].dirm |= 1u << (4 * e + T.comp);
                }
            }
        }
    } else {
        P.comp_order = 0;
    }
    return MM_OK;
}

static bool size_supported(int ng, int ne) {
#define MM_SUP(G, E) \
    if (ng == G && ne == E) return true;
    MM_SIZE_LIST(MM_SUP)
#undef MM_SUP
    return false;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int mm_version(void) { return 100; }

int mm_rate_plan_create(mm_rate_plan** out, int n_g, int n_e, int n_b, const double* E, const double* zd,
                        const double* d2, const double* dirs, const double* freqs, const int* pols,
                        const double* s0s, const int* comp, const double* split_sign, int n_comp, double Gamma,
                        double k, double Geff, double hbar) {
    mm_clear_error();
    if (!out) return mm_fail(MM_EINVAL, "plan: null output pointer");
    *out = nullptr;
    mm_rate_plan* pl = new (std::nothrow) mm_rate_plan;
    if (!pl) return mm_fail(MM_EINVAL, "plan: host allocation failed");
    pl->block = 0;
    pl->spec = pl->spec_grad = pl->spec_rf = nullptr;
    pl->spec_rf_peer = nullptr;
    pl->spec_rf_serial = 0;
    static std::atomic<unsigned long long> serials{0};
    pl->serial = ++serials;
    int rc = build_plan_host(pl->host, n_g, n_e, n_b, E, zd, d2, dirs, freqs, pols, s0s, comp, split_sign, n_comp,
                             Gamma, k, Geff, hbar);
    if (rc != MM_OK) {
        delete pl;
        return rc;
    }
    if (!size_supported(n_g, n_e)) {
        delete pl;
        return mm_fail(MM_EUNSUPPORTED, "plan: (n_ground, n_excited) = (%d, %d) has no compiled kernel", n_g, n_e);
    }
    pl->dev = nullptr;
    if (hipGetDevice(&pl->device) != hipSuccess) {
        delete pl;
        return mm_hip_fail("hipGetDevice");
    }
    if (hipMalloc((void**)&pl->dev, sizeof(MMRatePlanDev)) != hipSuccess) {
        delete pl;
        return mm_hip_fail("hipMalloc(plan)");
    }
    if (hipMemcpy(pl->dev, &pl->host, sizeof(MMRatePlanDev), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(pl->dev);
        delete pl;
        return mm_hip_fail("hipMemcpy(plan)");
    }
    *out = pl;
    return MM_OK;
}

long mm_rate_spec_source_tables(int n_g, int n_e, int n_b, const double* E, const double* zd, const double* d2,
                                const double* dirs, const double* freqs, const int* pols, const double* s0s,
                                const int* comp, const double* split_sign, int n_comp, double Gamma, double k,
                                double Geff, double hbar, int kind, char* buf, long size) {
    mm_clear_error();
    std::vector<MMRatePlanDev> P(1);  // ~80 KB: on the heap
    const int rc = build_plan_host(P[0], n_g, n_e, n_b, E, zd, d2, dirs, freqs, pols, s0s, comp, split_sign, n_comp,
                                   Gamma, k, Geff, hbar);
    if (rc != MM_OK) return rc;
    return mm_spec_source_copy(P[0], kind, buf, size);
}

int mm_rate_plan_destroy(mm_rate_plan* pl) {
    mm_clear_error();
    if (!pl) return MM_OK;
    int rc = MM_OK;
    mm_spec_release(pl);
    if (pl->dev && hipFree(pl->dev) != hipSuccess) rc = mm_hip_fail("hipFree(plan)");
    delete pl;
    return rc;
}

int mm_rate_plan_n_terms(const mm_rate_plan* pl) { return pl ? pl->host.n_terms : -1; }

int mm_rate_eq_map(const mm_rate_plan* pl, const double* v, int n_v, const double* z, int n_z, double B,
                   double* F, double* pop, double* R, void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "rate_eq_map: null plan");
    if (n_v < 0 || n_z < 0) return mm_fail(MM_EINVAL, "rate_eq_map: negative grid size");
    const long n = (long)n_v * (long)n_z;
    if (n == 0) return MM_OK;
    if (!v || !z || !F) return mm_fail(MM_EINVAL, "rate_eq_map: null pointer");
    hipStream_t s = (hipStream_t)stream;
    if (mm_spec_ready(pl)) return mm_spec_launch_map(pl, v, z, n_z, n, B, F, pop, R, s);
    const int ng = pl->host.n_g, ne = pl->host.n_e;
#define MM_MAP(G, E) \
    if (ng == G && ne == E) return launch_map<G, E>(pl, v, z, n_z, n, B, F, pop, R, s);
    MM_SIZE_LIST(MM_MAP)
#undef MM_MAP
    return mm_fail(MM_EUNSUPPORTED, "rate_eq_map: no kernel for (%d, %d)", ng, ne);
}

int mm_rate_eq_points(const mm_rate_plan* pl, long n, const double* v, const double* z, double B, double* F,
                      double* pop, double* R, void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "rate_eq_points: null plan");
    if (n > 0 && (!v || !z || !F)) return mm_fail(MM_EINVAL, "rate_eq_points: null pointer");
    if (n < 0) return mm_fail(MM_EINVAL, "rate_eq_points: negative size");
    if (n == 0) return MM_OK;
    hipStream_t s = (hipStream_t)stream;
    if (mm_spec_ready(pl)) return mm_spec_launch_points(pl, v, z, n, B, F, pop, R, s);
    const int ng = pl->host.n_g, ne = pl->host.n_e;
#define MM_PTS(G, E) \
    if (ng == G && ne == E) return launch_points<G, E>(pl, v, z, n, B, F, pop, R, s);
    MM_SIZE_LIST(MM_PTS)
#undef MM_PTS
    return mm_fail(MM_EUNSUPPORTED, "rate_eq_points: no kernel for (%d, %d)", ng, ne);
}

int mm_traj_euler_1d(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double mass,
                     double dt, int n_steps, double z_escape, double* zt, double* vt, double* zf, double* vf,
                     int* na, double kick_scale, unsigned int k0, unsigned int k1, unsigned int sid,
                     void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "traj_euler_1d: null plan");
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "traj_euler_1d: n >= 0 and n_steps >= 1 required");
    if (n == 0) return MM_OK;
    if (!z0 || !v0 || !zf || !vf || !na) return mm_fail(MM_EINVAL, "traj_euler_1d: null pointer");
    if ((zt == nullptr) != (vt == nullptr)) return mm_fail(MM_EINVAL, "traj_euler_1d: z_traj and v_traj go together");
    if (!(mass > 0.0)) return mm_fail(MM_EINVAL, "traj_euler_1d: mass must be positive");
    if (kick_scale < 0.0) return mm_fail(MM_EINVAL, "traj_euler_1d: kick_scale must be >= 0");
    hipStream_t s = (hipStream_t)stream;
    const int ng = pl->host.n_g, ne = pl->host.n_e;
#define MM_TRJ(G, E)                                                                                    \
    if (ng == G && ne == E)                                                                             \
        return launch_traj<G, E>(pl, n, z0, v0, B, mass, dt, n_steps, z_escape, zt, vt, zf, vf, na,     \
                                 kick_scale, k0, k1, sid, s);
    MM_SIZE_LIST(MM_TRJ)
#undef MM_TRJ
    return mm_fail(MM_EUNSUPPORTED, "traj_euler_1d: no kernel for (%d, %d)", ng, ne);
}

int mm_traj_euler_schedule(const mm_rate_plan* pl, long n, int n_steps, int save_full, double kick_scale) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "traj_euler_schedule: null plan");
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "traj_euler_schedule: n >= 0 and n_steps >= 1 required");
    const int ng = pl->host.n_g, ne = pl->host.n_e;
#define MM_TRJ(G, E) \
    if (ng == G && ne == E) return traj_auto_schedule<G, E>(pl, n, n_steps, save_full != 0, kick_scale);
    MM_SIZE_LIST(MM_TRJ)
#undef MM_TRJ
    return mm_fail(MM_EUNSUPPORTED, "traj_euler_schedule: no kernel for (%d, %d)", ng, ne);
}

int mm_traj_queue_set_wait(double seconds) {
    mm_clear_error();
    if (seconds != seconds || seconds > 86400.0)
        return mm_fail(MM_EINVAL, "traj_queue_set_wait: the bound must be a number of seconds up to one day");
    mm_queue_wait_override_us.store(seconds < 0.0 ? -1ll : (long long)(seconds * 1e6), std::memory_order_relaxed);
    return MM_OK;
}

long mm_traj_euler_workspace_bytes(long n, int n_steps) {
    MMQueueShape q;
    return mm_queue_shape(n, (long)n_steps - 1, q) ? q.bytes : 0;
}

int mm_traj_euler_1d_ws(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double mass,
                        double dt, int n_steps, double z_escape, double* zt, double* vt, double* zf, double* vf,
                        int* na, double kick_scale, unsigned int k0, unsigned int k1, unsigned int sid, int schedule,
                        void* workspace, long workspace_bytes, void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "traj_euler_1d_ws: null plan");
    if (schedule < 0 || schedule > 2)
        return mm_fail(MM_EINVAL, "traj_euler_1d_ws: schedule is 0 (auto), 1 (grid) or 2 (queue)");
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "traj_euler_1d_ws: n >= 0 and n_steps >= 1 required");
    if (n == 0) return MM_OK;
    if (!z0 || !v0 || !zf || !vf || !na) return mm_fail(MM_EINVAL, "traj_euler_1d_ws: null pointer");
    if ((zt == nullptr) != (vt == nullptr))
        return mm_fail(MM_EINVAL, "traj_euler_1d_ws: z_traj and v_traj go together");
    if (!(mass > 0.0)) return mm_fail(MM_EINVAL, "traj_euler_1d_ws: mass must be positive");
    if (kick_scale < 0.0) return mm_fail(MM_EINVAL, "traj_euler_1d_ws: kick_scale must be >= 0");
    if (workspace_bytes < 0) return mm_fail(MM_EINVAL, "traj_euler_1d_ws: workspace_bytes >= 0 required");
    if (workspace && ((uintptr_t)workspace & 255u))
        return mm_fail(MM_EINVAL, "traj_euler_1d_ws: the workspace must be 256-byte aligned");
    hipStream_t s = (hipStream_t)stream;
    const int ng = pl->host.n_g, ne = pl->host.n_e;
#define MM_TRJ(G, E)                                                                                       \
    if (ng == G && ne == E)                                                                                \
        return launch_traj_sched<G, E>(pl, n, z0, v0, B, mass, dt, n_steps, z_escape, zt, vt, zf, vf, na,   \
                                       kick_scale, k0, k1, sid, schedule, workspace, workspace_bytes, s);
    MM_SIZE_LIST(MM_TRJ)
#undef MM_TRJ
    return mm_fail(MM_EUNSUPPORTED, "traj_euler_1d_ws: no kernel for (%d, %d)", ng, ne);
}

// ---------------------------------------------------------------------------
// The same ensemble without a caller workspace, for callers that cannot
// allocate device memory: the queue lives in a file-scope device array of
// MM_NOALLOC_SCRATCH_BYTES.  Host code never passes the array's name to a
// kernel (in host code it is a stand-in, not a device address): its device
// address is looked up once per device, with that device current, and kept.
// ---------------------------------------------------------------------------
#define MM_NOALLOC_SCRATCH_BYTES (4l << 20)
static __device__ __align__(256) unsigned char scratch[MM_NOALLOC_SCRATCH_BYTES];

static void* g_scratch_ptr[64];
static hipError_t g_scratch_err[64];
static std::once_flag g_scratch_once[64];

static int scratch_address(void** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return mm_fail(MM_EHIP, "traj_euler_1d_noalloc: no current device");
    std::call_once(g_scratch_once[dev], [dev]() {
        void* p = nullptr;
        g_scratch_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[dev] = (g_scratch_err[dev] == hipSuccess) ? p : nullptr;
    });
    if (g_scratch_err[dev] != hipSuccess || g_scratch_ptr[dev] == nullptr)
        return mm_fail(MM_EHIP, "traj_euler_1d_noalloc: hipGetSymbolAddress(scratch): %s",
                       hipGetErrorString(g_scratch_err[dev]));
    *out = g_scratch_ptr[dev];
    return MM_OK;
}

long mm_traj_euler_noalloc_max_bytes(void) { return MM_NOALLOC_SCRATCH_BYTES; }

int mm_traj_euler_1d_noalloc(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B,
                             double mass, double dt, int n_steps, double z_escape, double* zf, double* vf, int* na,
                             double kick_scale, unsigned int k0, unsigned int k1, unsigned int sid, int schedule,
                             void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "traj_euler_1d_noalloc: null plan");
    if (schedule < 0 || schedule > 2)
        return mm_fail(MM_EINVAL, "traj_euler_1d_noalloc: schedule is 0 (auto), 1 (grid) or 2 (queue)");
    if (n < 0 || n_steps < 1)
        return mm_fail(MM_EINVAL, "traj_euler_1d_noalloc: n >= 0 and n_steps >= 1 required");
    if (n == 0) return MM_OK;
    MMQueueShape q;
    const bool queue = schedule != 1 && mm_queue_shape(n, (long)n_steps - 1, q);
    if (queue && q.bytes > MM_NOALLOC_SCRATCH_BYTES)
        return mm_fail(MM_EINVAL,
                       "traj_euler_1d_noalloc: %ld particles x %d steps need %ld bytes of queue, more than the "
                       "%ld-byte scratch array (use schedule 1, or mm_traj_euler_1d_ws with a workspace)",
                       n, n_steps, q.bytes, (long)MM_NOALLOC_SCRATCH_BYTES);
    void* ws = nullptr;
    if (queue) {
        const int rc = scratch_address(&ws);
        if (rc != MM_OK) return rc;
    }
    // the queue kernels clear the part of the scratch array they use with a
    // kernel before every launch
    const int rc = mm_traj_euler_1d_ws(pl, n, z0, v0, B, mass, dt, n_steps, z_escape, nullptr, nullptr, zf, vf, na,
                                       kick_scale, k0, k1, sid, schedule, ws, queue ? MM_NOALLOC_SCRATCH_BYTES : 0,
                                       stream);
    if (rc != MM_OK) {
        // keep the message but name this entry point
        std::string m = mm_err_buf;
        return mm_fail(rc, "traj_euler_1d_noalloc: %s", m.c_str());
    }
    return MM_OK;
}

long mm_traj_stats_workspace_bytes(long n, int n_samp) {
    if (n < 0 || n_samp < 0) return -1;
    return ((n + 63) / 64) * (long)n_samp * 5 * 8;
}

int mm_traj_stats_1d(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double mass,
                     double dt, int n_steps, double z_escape, double r_alive, const int* sample_steps, int n_samp,
                     double* stats, double kick_scale, unsigned int k0, unsigned int k1, unsigned int sid,
                     void* workspace, long workspace_bytes, void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "traj_stats_1d: null plan");
    if (n < 1 || n_steps < 1 || n_samp < 1)
        return mm_fail(MM_EINVAL, "traj_stats_1d: n >= 1, n_steps >= 1 and n_samp >= 1 required");
    if (!z0 || !v0 || !sample_steps || !stats || !workspace)
        return mm_fail(MM_EINVAL, "traj_stats_1d: null pointer");
    if (!(mass > 0.0)) return mm_fail(MM_EINVAL, "traj_stats_1d: mass must be positive");
    if (kick_scale < 0.0) return mm_fail(MM_EINVAL, "traj_stats_1d: kick_scale must be >= 0");
    if (workspace_bytes < mm_traj_stats_workspace_bytes(n, n_samp))
        return mm_fail(MM_EINVAL, "traj_stats_1d: workspace smaller than mm_traj_stats_workspace_bytes()");
    const int ng = pl->host.n_g, ne = pl->host.n_e;
#define MM_TST(G, E)                                                                                       \
    if (ng == G && ne == E)                                                                                \
        return launch_traj_stats<G, E>(pl, n, z0, v0, B, mass, dt, n_steps, z_escape, r_alive, sample_steps, \
                                       n_samp, (double*)workspace, stats, kick_scale, k0, k1, sid,          \
                                       (hipStream_t)stream);
    MM_SIZE_LIST(MM_TST)
#undef MM_TST
    return mm_fail(MM_EUNSUPPORTED, "traj_stats_1d: no kernel for (%d, %d)", ng, ne);
}

int mm_traj_euler_rf_1d(const mm_rate_plan* p0, const mm_rate_plan* p1, long n, const double* z0,
                        const double* v0, double B, double mass, double dt, int n_steps, double z_escape,
                        double* zt, double* vt, double* zf, double* vf, int* na, double kick_scale, unsigned int k0,
                        unsigned int k1, unsigned int sid, void* stream) {
    mm_clear_error();
    if (!p0 || !p1) return mm_fail(MM_EINVAL, "traj_euler_rf_1d: null plan");
    if (p0->host.n_g != p1->host.n_g || p0->host.n_e != p1->host.n_e)
        return mm_fail(MM_EINVAL, "traj_euler_rf_1d: the two phase plans differ in size");
    if (!p0->host.unit_dirs || !p1->host.unit_dirs)
        return mm_fail(MM_EUNSUPPORTED, "traj_euler_rf_1d: beams must be along +-z");
    if (n < 0 || n_steps < 1) return mm_fail(MM_EINVAL, "traj_euler_rf_1d: n >= 0 and n_steps >= 1 required");
    if (n == 0) return MM_OK;
    if (!z0 || !v0 || !zf || !vf || !na) return mm_fail(MM_EINVAL, "traj_euler_rf_1d: null pointer");
    if ((zt == nullptr) != (vt == nullptr))
        return mm_fail(MM_EINVAL, "traj_euler_rf_1d: z_traj and v_traj go together");
    if (!(mass > 0.0)) return mm_fail(MM_EINVAL, "traj_euler_rf_1d: mass must be positive");
    if (kick_scale < 0.0) return mm_fail(MM_EINVAL, "traj_euler_rf_1d: kick_scale must be >= 0");
    hipStream_t s = (hipStream_t)stream;
    const int ng = p0->host.n_g, ne = p0->host.n_e;
#define MM_TRJRF(G, E)                                                                                  \
    if (ng == G && ne == E)                                                                             \
        return launch_traj_rf<G, E>(p0, p1, n, z0, v0, B, mass, dt, n_steps, z_escape, zt, vt, zf, vf, na, \
                                    kick_scale, k0, k1, sid, s);
    MM_SIZE_LIST(MM_TRJRF)
#undef MM_TRJRF
    return mm_fail(MM_EUNSUPPORTED, "traj_euler_rf_1d: no kernel for (%d, %d)", ng, ne);
}

int mm_philox4x32_10(long n, const unsigned int* ctr, unsigned int k0, unsigned int k1, unsigned int* out,
                     void* stream) {
    mm_clear_error();
    if (n < 0) return mm_fail(MM_EINVAL, "philox: negative size");
    if (n == 0) return MM_OK;
    if (!ctr || !out) return mm_fail(MM_EINVAL, "philox: null pointer");
    hipLaunchKernelGGL(k_philox, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream, n, ctr, k0, k1, out);
    return mm_check_launch("k_philox");
}

}  // extern "C"

extern "C" const char* mm_last_error(void) { return mm_err_buf; }
