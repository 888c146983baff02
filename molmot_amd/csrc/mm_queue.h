// Work queue of the persistent trajectory kernels (the 1D ensemble, its
// forward sensitivities, its Monte-Carlo temperature statistics and the
// per-axis 3D ensemble).
//
// One thread per particle and 64 per wavefront leave 1e5 particles as 1,563
// wavefronts on the 1,024 SIMDs of a 256-CU device: the SIMDs that carry two
// set the time while the others run one.  With the queue, a unit of work is
// one group of 64 particles over one chunk of steps; a fixed grid of
// wavefronts (w per SIMD) takes units from the queue, so a group moves
// between SIMDs from chunk to chunk and the load evens out.  A kernel keeps a
// particle's state in its output arrays between chunks.
//
// Layout (in the caller's workspace, 256-byte aligned): ctl[0] counts claimed
// units, ctl[1] published ones, ctl[2] flags a wait that timed out; from byte
// 256, item[p] is the p-th published unit, (chunk << 32) | group with chunk
// >= 1, 0 until written.  Units 0 .. groups-1 are chunk 0 of every group and
// need no item; unit u >= groups is item[u - groups], published by the
// wavefront that finished the group's previous chunk, after its state stores
// (release fence / acquire fence at agent scope).  Each item is written once,
// so no slot is reused.  A wavefront only waits for an item that a running
// wavefront will publish, so the queue cannot deadlock; the wait is still
// bounded (wait_ticks of the wall clock) so that every wavefront exits
// whatever happens, and k_mm_queue_status then marks every particle -1.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "mm_errors.h"
#include "mm_queue_dev.h"

// at most 256 chunks per trajectory (the queue holds groups x (chunks - 1)
// items of 8 bytes) and at least 128 steps per chunk.  Finer chunks even
// out the end of the run: at 1e5 particles x 30,000 steps, 64 chunks gave
// 1.40e10 particle-steps/s, 128 chunks 1.43e10, 235 (128 steps each)
// 1.45e10
#ifndef MM_TQ_MAX_CHUNKS
#define MM_TQ_MAX_CHUNKS 256
#endif
#define MM_TQ_MIN_CHUNK 128
// a wait for a queue item gives up after max(MM_TQ_WAIT_S, chunk steps x
// MM_TQ_WAIT_STEP_S) seconds (never in a correct run: an item is published
// within one chunk of steps, and a step of the heaviest plan takes ~0.1 ms
// of one wavefront's time)
#define MM_TQ_WAIT_S 10.0
#define MM_TQ_WAIT_STEP_S 0.01

// host override of that bound in microseconds (< 0: the rule above), set by
// mm_traj_queue_set_wait for every later launch of the process
inline std::atomic<long long> mm_queue_wait_override_us{-1};

// clears the queue (a kernel, not a memset, so that the launch sequence can
// be captured into a HIP graph)
static __global__ void k_mm_queue_clear(unsigned long long* __restrict__ w, long n_words) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_words) w[i] = 0ull;
}

// after a queue kernel: a timed-out wait leaves cnt = -1 for every particle
static __global__ void k_mm_queue_status(const unsigned* __restrict__ ctl, int* __restrict__ cnt, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && ctl[2] != 0u) cnt[i] = -1;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct MMQueueShape {
    unsigned groups, chunks;
    int chunk;
    long bytes;
};

// n particles over `steps` updates: false if the queue does not apply
static inline bool mm_queue_shape(long n, long steps, MMQueueShape& q) {
    if (n <= 0 || steps < 2 * MM_TQ_MIN_CHUNK) return false;
    const long groups = (n + 63) / 64;
    long L = (steps + MM_TQ_MAX_CHUNKS - 1) / MM_TQ_MAX_CHUNKS;
    if (L < MM_TQ_MIN_CHUNK) L = MM_TQ_MIN_CHUNK;
    L += L & 1;  // chunks start at odd steps of the 1D ensemble (Philox pairs)
    const long chunks = (steps + L - 1) / L;
    if (groups * chunks >= (1l << 31)) return false;
    q.groups = (unsigned)groups;
    q.chunks = (unsigned)chunks;
    q.chunk = (int)L;
    q.bytes = 256 + 8 * groups * (chunks - 1);
    return true;
}

struct MMQueueDev {
    int ok, n_cu, lds_cu, lds_block;
    unsigned long long ticks_per_s;
};

// per-device attributes, queried once per device (std::call_once: first
// calls from several host threads are safe)
inline MMQueueDev mm_queue_dev_table[64];
inline std::once_flag mm_queue_dev_once[64];

static inline int mm_queue_device(const MMQueueDev** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return mm_fail(MM_EHIP, "trajectory queue: no current device");
    MMQueueDev& d = mm_queue_dev_table[dev];
    std::call_once(mm_queue_dev_once[dev], [&d, dev]() {
        int khz = 0;
        d.ok = hipDeviceGetAttribute(&d.n_cu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
               hipDeviceGetAttribute(&d.lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) ==
                   hipSuccess &&
               hipDeviceGetAttribute(&d.lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
               hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && d.n_cu > 0 &&
               khz > 0;
        d.ticks_per_s = (unsigned long long)(khz > 0 ? khz : 0) * 1000ull;
    });
    if (!d.ok) return mm_fail(MM_EHIP, "trajectory queue: device attributes unavailable");
    *out = &d;
    return MM_OK;
}

// Where a queue kernel would run: the device, its SIMDs, the kernel's
// resident 64-thread blocks per CU and on the device, its static LDS.
struct MMQueueSlots {
    const MMQueueDev* d;
    long n_simd, resident, lds_static;
    int per_cu;
};

static inline int mm_queue_slots(const void* kern, MMQueueSlots& s) {
    const int rc = mm_queue_device(&s.d);
    if (rc != MM_OK) return rc;
    hipFuncAttributes fa;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&s.per_cu, kern, 64, 0) != hipSuccess || s.per_cu <= 0 ||
        hipFuncGetAttributes(&fa, kern) != hipSuccess)
        return mm_fail(MM_EHIP, "trajectory queue: occupancy query failed");
    s.n_simd = 4l * s.d->n_cu;
    s.resident = (long)s.per_cu * s.d->n_cu;
    s.lds_static = (long)fa.sharedSizeBytes;
    return MM_OK;
}

// The same for a kernel of a run-time compiled module.
static inline int mm_queue_slots_fn(hipFunction_t f, MMQueueSlots& s) {
    const int rc = mm_queue_device(&s.d);
    if (rc != MM_OK) return rc;
    int lds = 0;
    if (hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&s.per_cu, f, 64, 0) != hipSuccess || s.per_cu <= 0 ||
        hipFuncGetAttribute(&lds, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, f) != hipSuccess)
        return mm_fail(MM_EHIP, "trajectory queue: occupancy query failed (specialised kernel)");
    s.n_simd = 4l * s.d->n_cu;
    s.resident = (long)s.per_cu * s.d->n_cu;
    s.lds_static = (long)lds;
    return MM_OK;
}

// The auto rule.  The queue gains nothing where every SIMD is busy anyway
// (1D ensemble at 2,048 groups: 1.64e10 against 1.66e10 particle-steps/s;
// its step loop carries more SGPR spills), and gains where one thread per
// particle leaves a last round of wavefronts well short of full (1e5
// particles: 1,563 groups, 1.45e10 against 1.24e10).  So: the grid up to one wavefront per SIMD, and wherever
// the last round -- over the SIMDs while all groups fit at once, else over
// all resident slots -- is at least 90% full.
static inline bool mm_queue_pays(const MMQueueShape& q, const MMQueueSlots& s) {
    const long G = (long)q.groups;
    if (G <= s.n_simd) return false;
    if (G <= s.resident) {
        const long w_all = (G + s.n_simd - 1) / s.n_simd;
        return 10l * G < 9l * w_all * s.n_simd;
    }
    const long rem = G % s.resident;
    return rem != 0 && 10l * rem < 9l * s.resident;
}

// Grid of the persistent kernel: w wavefronts per SIMD, enough for every
// group and at most what fits; each block's LDS padded to 1/(4w) of the
// CU's so that no SIMD holds more than w of them.
static inline void mm_queue_grid(const MMQueueShape& q, const MMQueueSlots& s, unsigned* grid, size_t* dyn) {
    long w = ((long)q.groups + s.n_simd - 1) / s.n_simd;
    if (w > s.per_cu / 4) w = s.per_cu / 4;
    if (w < 1) w = 1;
    long pad = (long)s.d->lds_cu / (4 * w) - s.lds_static;
    if (pad > (long)s.d->lds_block - s.lds_static) pad = (long)s.d->lds_block - s.lds_static;
    pad &= ~511l;
    if (pad < 0) pad = 0;
    *grid = (unsigned)(w * s.n_simd);
    *dyn = (size_t)pad;
}

static inline MMQueue mm_queue_make(void* ws, const MMQueueShape& q, const MMQueueSlots& s) {
    MMQueue Q;
    Q.ctl = (unsigned*)ws;
    Q.item = (unsigned long long*)((char*)ws + 256);
    Q.groups = q.groups;
    Q.chunks = q.chunks;
    Q.chunk = q.chunk;
    const long long us = mm_queue_wait_override_us.load(std::memory_order_relaxed);
    double sec = MM_TQ_WAIT_STEP_S * (double)q.chunk;
    if (sec < MM_TQ_WAIT_S) sec = MM_TQ_WAIT_S;
    if (us >= 0) sec = 1e-6 * (double)us;
    Q.wait_ticks = (unsigned long long)(sec * (double)s.d->ticks_per_s);
    return Q;
}
