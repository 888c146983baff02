// Plan-specialised trajectory kernels (rate_spec.hip): the interface the
// trajectory launchers of rate_eq.hip use.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mm_plan_host.h"
#include "mm_queue.h"

// The generated source for a plan's host tables into buf (NUL-terminated,
// truncated to size); returns the size it needs, or a negative MM_* code
// (with the error message set) if the plan cannot be specialised.
long mm_spec_source_copy(const MMRatePlanDev& P, char* buf, long size);

// Releases the plan's reference to its specialised module (no-op if none).
void mm_spec_release(mm_rate_plan* pl);

// True when the plan carries specialised kernels for the current device.
bool mm_spec_ready(const mm_rate_plan* pl);

// Occupancy of the specialised queue kernel (kicks or not), for the queue's
// auto rule and grid.
int mm_spec_queue_slots(const mm_rate_plan* pl, bool kicks, MMQueueSlots& s);

// One-thread-per-particle launch (finals only) of the specialised kernel.
int mm_spec_launch_grid(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                        double dt, int ns, double ze, double* zf, double* vf, int* na, double ks, uint32_t k0,
                        uint32_t k1, uint32_t sid, hipStream_t s);

// Work-queue launch of the specialised kernel (the caller clears the queue
// before and runs the status kernel after, as for the generic kernel).
int mm_spec_launch_queue(const mm_rate_plan* pl, unsigned grid, size_t dyn, long n, const double* z0,
                         const double* v0, double B, double m, double dt, int ns, double ze, double* zf, double* vf,
                         int* na, double ks, uint32_t k0, uint32_t k1, uint32_t sid, const MMQueue& Q,
                         hipStream_t s);
