// Device side of the trajectory work queue (layout and protocol: mm_queue.h):
// the queue descriptor, claiming a unit, publishing a group's next chunk.
// Free of host headers, so that run-time compiled kernels include it too.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

struct MMQueue {
    unsigned* ctl;
    unsigned long long* item;
    unsigned groups, chunks;
    int chunk;  // steps per chunk (even)
    unsigned long long wait_ticks;
};

// Claims the next unit for the calling wavefront (group g, chunk c).  False
// once every unit is taken or a wait has timed out: the wavefront exits.
__device__ __forceinline__ bool mm_queue_next(const MMQueue& Q, unsigned& g, unsigned& c) {
    unsigned u = 0;
    if (threadIdx.x == 0) u = __hip_atomic_fetch_add(&Q.ctl[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    u = (unsigned)__builtin_amdgcn_readfirstlane((int)u);
    if (u >= Q.groups * Q.chunks) return false;
    if (__hip_atomic_load(&Q.ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) return false;
    g = u;
    c = 0;
    if (u < Q.groups) return true;
    const unsigned long long t0 = wall_clock64();
    // poll without acquire (an acquire invalidates the L2 of this XCD for
    // every wavefront on it); one acquire fence once the item is there
    for (;;) {
        const unsigned long long e =
            __hip_atomic_load(&Q.item[u - Q.groups], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (e != 0ull) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            g = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)e);
            c = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(e >> 32));
            return true;
        }
        if (wall_clock64() - t0 > Q.wait_ticks ||
            __hip_atomic_load(&Q.ctl[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
            if (threadIdx.x == 0) __hip_atomic_store(&Q.ctl[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(32);  // ~1 us: pickup latency is immaterial
    }
}

// Hands group g on to its next chunk; call after every lane's state stores.
__device__ __forceinline__ void mm_queue_publish(const MMQueue& Q, unsigned g, unsigned c) {
    if (c + 1 >= Q.chunks) return;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (threadIdx.x == 0) {
        const unsigned p = __hip_atomic_fetch_add(&Q.ctl[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (p < Q.groups * (Q.chunks - 1))  // ordered after the state by the fence above
            __hip_atomic_store(&Q.item[p], ((unsigned long long)(c + 1) << 32) | g, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        else
            __hip_atomic_store(&Q.ctl[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
