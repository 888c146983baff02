// Plan-specialised trajectory kernels (rate_spec.hip): the interface the
// trajectory launchers of rate_eq.hip use.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mm_plan_host.h"
#include "mm_queue.h"
#include "mm_vjp.h"

// kinds of specialised kernels (bit mask of mm_rate_plan_specialize)
#define MM_SPEC_FORWARD 1
#define MM_SPEC_GRADIENT 2
// the RF MOT's pair of phase plans (mm_rate_plan_specialize_rf; not in the mask)
#define MM_SPEC_RF 4

// The generated source of one kind for a plan's host tables into buf
// (NUL-terminated, truncated to size); returns the size it needs, or a
// negative MM_* code (with the error message set) if the plan cannot have
// kernels of that kind.
long mm_spec_source_copy(const MMRatePlanDev& P, int kind, char* buf, long size);

// Releases the plan's reference to its specialised module (no-op if none).
void mm_spec_release(mm_rate_plan* pl);

// True when the plan carries specialised forward (gradient) kernels for the
// current device.
bool mm_spec_ready(const mm_rate_plan* pl);
bool mm_spec_grad_ready(const mm_rate_plan* pl);

// Occupancy of the specialised queue kernel (kicks or not), for the queue's
// auto rule and grid.
int mm_spec_queue_slots(const mm_rate_plan* pl, bool kicks, MMQueueSlots& s);

// One-thread-per-particle launch (finals only) of the specialised kernel.
int mm_spec_launch_grid(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                        double dt, int ns, double ze, double* zf, double* vf, int* na, double ks, uint32_t k0,
                        uint32_t k1, uint32_t sid, hipStream_t s);

// The same with the trajectory saved (zt, vt: n x ns).
int mm_spec_launch_grid_save(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                             double dt, int ns, double ze, double* zt, double* vt, double* zf, double* vf, int* na,
                             double ks, uint32_t k0, uint32_t k1, uint32_t sid, hipStream_t s);

// Work-queue launch of the specialised kernel (the caller clears the queue
// before and runs the status kernel after, as for the generic kernel).
int mm_spec_launch_queue(const mm_rate_plan* pl, unsigned grid, size_t dyn, long n, const double* z0,
                         const double* v0, double B, double m, double dt, int ns, double ze, double* zf, double* vf,
                         int* na, double ks, uint32_t k0, uint32_t k1, uint32_t sid, const MMQueue& Q,
                         hipStream_t s);

// The forward-sensitivity kernels (rate_adjoint.hip's contracts).
int mm_spec_tangent_slots(const mm_rate_plan* pl, MMQueueSlots& s);
int mm_spec_launch_tangent_grid(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B,
                                double m, double dt, int ns, const int* n_up, double ze, double* zf, double* vf,
                                int* nd, double* sens, hipStream_t s);
int mm_spec_launch_tangent_queue(const mm_rate_plan* pl, unsigned grid, size_t dyn, long n, const double* z0,
                                 const double* v0, double B, double m, double dt, int ns, const int* n_up, double ze,
                                 double* zf, double* vf, int* nd, double* sens, const MMQueue& Q, hipStream_t s);

// The checkpointed adjoint's forward pass (forward module) and segment
// (gradient module), VjpArgs of mm_vjp.h.
int mm_spec_launch_vjp_forward(const mm_rate_plan* pl, const VjpArgs& A, hipStream_t s);
int mm_spec_launch_vjp_segment(const mm_rate_plan* pl, const VjpArgs& A, int seg, hipStream_t s);

// Force map and scattered points (forward module: rate_eq.hip's contracts)
// and the force gradient at scattered points (gradient module).
int mm_spec_launch_map(const mm_rate_plan* pl, const double* v, const double* z, int nz, long n, double B, double* F,
                       double* pop, double* R, hipStream_t s);
int mm_spec_launch_points(const mm_rate_plan* pl, const double* v, const double* z, long n, double B, double* F,
                          double* pop, double* R, hipStream_t s);
int mm_spec_launch_grad_points(const mm_rate_plan* pl, const double* v, const double* z, long n, double B, double* F,
                               double* J, hipStream_t s);

// Monte-Carlo temperature statistics (forward module; traj3d.hip's
// contracts of k_traj_vsq without RF and k_traj_vsq_queue).
int mm_spec_vsq_slots(const mm_rate_plan* pl, MMQueueSlots& s);
int mm_spec_launch_vsq_grid(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                            double dt, int ns, int ieq, double ze, double* vsq, long long* cnt, hipStream_t s);
int mm_spec_launch_vsq_queue(const mm_rate_plan* pl, unsigned grid, size_t dyn, long n, const double* z0,
                             const double* v0, double B, double m, double dt, int ns, int ieq, double ze,
                             double* vsq, long long* cnt, double* zs, double* vs, const MMQueue& Q, hipStream_t s);

// Per-axis 3D model (forward module; traj3d.hip's contracts of
// k_traj_axis3d without saved trajectory or RF, and k_traj_axis3d_queue).
int mm_spec_axis3d_slots(const mm_rate_plan* pl, MMQueueSlots& s);
int mm_spec_launch_axis3d_grid(const mm_rate_plan* pl, long n, const double* r0, const double* v0, double B, double m,
                               double dt, int ns, double re, double* rf, double* vf, int* na, hipStream_t s);
int mm_spec_launch_axis3d_queue(const mm_rate_plan* pl, unsigned grid, size_t dyn, long n, const double* r0,
                                const double* v0, double B, double m, double dt, int ns, double re, double* rf,
                                double* vf, int* na, const MMQueue& Q, hipStream_t s);

// Ensemble statistics at sample steps (forward module; rate_eq.hip's
// contract of k_traj_stats: one partial per 64 particles in part).
int mm_spec_launch_stats(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                         double dt, int ns, double ze, double ra, const int* samp, int n_samp, double* part, double ks,
                         uint32_t k0, uint32_t k1, uint32_t sid, hipStream_t s);

// True when p0 carries the RF kernels built for the phase plans (p0, p1) on
// the current device; the launches below then run them (same contracts as
// k_traj_euler and k_traj_vsq with RF; zt == nullptr: no saved trajectory).
bool mm_spec_rf_ready(const mm_rate_plan* p0, const mm_rate_plan* p1);
int mm_spec_launch_rf_traj(const mm_rate_plan* p0, const mm_rate_plan* p1, long n, const double* z0,
                           const double* v0, double B, double m, double dt, int ns, double ze, double* zt, double* vt,
                           double* zf, double* vf, int* na, double ks, uint32_t k0, uint32_t k1, uint32_t sid,
                           hipStream_t s);
int mm_spec_launch_rf_vsq(const mm_rate_plan* p0, const mm_rate_plan* p1, long n, const double* z0,
                          const double* v0, double B, double m, double dt, int ns, int ieq, double ze, double* vsq,
                          long long* cnt, hipStream_t s);
