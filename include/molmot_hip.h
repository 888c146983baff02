/*
 * molmot_hip -- C ABI of the MI355X (gfx950) rate-equation / trajectory
 * kernels.  Plain pointers and sizes only.
 *
 * Conventions
 *   - "host" pointers are CPU memory, "device" pointers are HIP device memory
 *     of the current device (e.g. torch.Tensor.data_ptr() of a cuda tensor).
 *   - `stream` is a hipStream_t (NULL = default stream).  Calls enqueue work
 *     and return without synchronising; they allocate nothing, so they may be
 *     captured into a hipGraph.
 *   - Every function returns MM_OK (0) or a negative MM_E* code and never
 *     throws; mm_last_error() describes the last failure of the calling
 *     thread.
 *   - All arrays are FP64, C-contiguous.  Polarisation indices are 0,1,2 for
 *     q = -1, 0, +1 (sigma-, pi, sigma+).
 */
#ifndef MOLMOT_HIP_H
#define MOLMOT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MM_OK 0
#define MM_EINVAL (-1)       /* bad argument / size */
#define MM_EUNSUPPORTED (-2) /* (n_ground, n_excited) not instantiated */
#define MM_EHIP (-3)         /* HIP runtime error */

typedef struct mm_rate_plan mm_rate_plan;

/* Last error message of the calling thread ("" if none). */
const char* mm_last_error(void);

/* ABI version (major*100 + minor). */
int mm_version(void);

/*
 * Build a device-resident plan for one molecule and beam set.
 * Replaces CUDAMolData.__init__ (molmot_cuda/rate_equations_cuda.py:399-441)
 * together with the per-call beam packing of solve_rate_equations_jit
 * (molmot/obe/rate_equations_jit.py:155-159).
 *   energies[n_s], zeeman_z_diag[n_s]          n_s = n_ground + n_excited (host)
 *   d_squared[n_ground][n_excited][3]           |tdm|^2 (host)
 *   beam_dirs/freqs/s0s[n_beams], beam_pols[n_beams]           (host)
 *   beam_comp[n_beams] (frequency component, may be NULL),
 *   comp_split_sign[n_comp] (d freq / d split per component, may be NULL)
 *   hbar: the constant the force uses (the reference JIT path hard-codes
 *   1.0545718e-34, molmot/obe/rate_equations_jit.py:27).
 * Allocates device memory on the current device; it and mm_rate_plan_destroy
 * are the only calls that allocate or free.
 */
int mm_rate_plan_create(mm_rate_plan** plan, int n_ground, int n_excited, int n_beams,
                        const double* energies, const double* zeeman_z_diag,
                        const double* d_squared, const double* beam_dirs,
                        const double* beam_freqs, const int* beam_pols,
                        const double* beam_s0s, const int* beam_comp,
                        const double* comp_split_sign, int n_comp,
                        double Gamma, double k_wave, double Gamma_eff, double hbar);

/* Release a plan (NULL is accepted). */
int mm_rate_plan_destroy(mm_rate_plan* plan);

/* Number of (ground, excited, beam) coupling terms kept by the plan. */
int mm_rate_plan_n_terms(const mm_rate_plan* plan);

/*
 * Threads per block of the map / point kernels: 64, 128, 256 (the per-thread
 * LDS slab of 2 n_ground doubles must fit 48 KiB), or 0 for the default (the
 * largest that fits).  A tuning knob (molmot_amd.autotune); the reference's
 * block_size argument (molmot_cuda/rate_equations_cuda.py:448-544) is the
 * same choice made by hand.  get_block returns the value in effect.
 */
int mm_rate_plan_set_block(mm_rate_plan* plan, int block);
int mm_rate_plan_get_block(const mm_rate_plan* plan);

/*
 * Plan-specialised trajectory kernels.  The reference compiles its GPU
 * kernels at run time from source strings with the molecule's tables as
 * kernel arguments (CuPy RawKernel, molmot_cuda/trajectories_cuda.py:44-300,
 * built per CUDAMolData, molmot_cuda/rate_equations_cuda.py:399-441); here
 * mm_rate_plan_specialize generates kernels with the plan's tables as
 * constants of the code (straight-line evaluation over the non-empty
 * (ground, excited) pairs) and compiles them with hipRTC for the plan's
 * device; modules are cached per process by generated source.  kinds is a
 * bit mask:
 *   1 (forward): the 1D trajectory and force kernels (~1 s to build).
 *     Later trajectory launches on the plan (mm_traj_euler_1d and
 *     mm_traj_euler_1d_ws, with or without the saved trajectory,
 *     mm_traj_euler_1d_noalloc), mm_rate_eq_map, mm_rate_eq_points,
 *     mm_traj_vsq_1d / _ws (no RF), mm_traj_axis3d / _ws (no saved
 *     trajectory, no RF), mm_traj_stats_1d and the forward pass of
 *     mm_traj_vjp_1d run them,
 *     with results bit-identical to the generic kernels'.
 *   2 (gradient): the force-gradient kernels (a few seconds to build);
 *     mm_traj_tangent_1d, mm_traj_tangent_1d_ws, the segments of
 *     mm_traj_vjp_1d and mm_rate_eq_grad_points run them.  Their force
 *     gradient is the generic kernels' reverse pass rearranged (the adjoint
 *     of each ground state's total rate is completed before its pairs are
 *     revisited), equal to it to rounding.
 *   0 releases every specialised kernel of the plan.
 * Applies to packed plans (unit beam directions, n_excited <= 4, every pair
 * empty or driven by two +z and two -z beams) with no dark ground state,
 * and for kind 2 one beam of each of at most four frequency components per
 * pair; else MM_EUNSUPPORTED.  Not thread-safe against concurrent launches
 * on the same plan.
 * mm_rate_plan_is_specialized: the mask of kinds the plan carries.
 * mm_rate_plan_spec_source: the generated source of one kind
 * (NUL-terminated, truncated to size) into buf; returns the size it needs,
 * or a negative MM_* code if the plan cannot have that kind.  buf may be
 * NULL.
 */
int mm_rate_plan_specialize(mm_rate_plan* plan, int kinds);
int mm_rate_plan_is_specialized(const mm_rate_plan* plan);
long mm_rate_plan_spec_source(const mm_rate_plan* plan, int kind, char* buf, long size);

/*
 * Plan-specialised RF MOT kernels for the two switching phases
 * (molmot/mot/simulator.py:49-90, whose step averages the force of the
 * phase-0 beams at +B' and of the phase-1 beams at -B'): p0 keeps kernels
 * with both plans' tables as constants, which mm_traj_euler_rf_1d and
 * mm_traj_vsq_rf_1d on (p0, p1) then run (bit-identical results).  RF phase
 * plans drive a pair with one +z and one -z beam per frequency, outside the
 * packed form; the generated pair sums follow the general path's order.
 * p1 == NULL releases them.  Both plans need unit beam directions, the same
 * state counts and no dark ground state (else MM_EUNSUPPORTED).  The kernels
 * belong to (p0, p1) as built: another second plan, or a plan built later at
 * p1's address, takes the generic kernels.  No reference counterpart (the
 * reference compiles one generic kernel, rate_equations_jit.py).
 * mm_rate_plan_is_specialized_rf: 1 when (p0, p1) has them;
 * mm_rate_plan_spec_source_rf: their generated source, as
 * mm_rate_plan_spec_source.
 */
int mm_rate_plan_specialize_rf(mm_rate_plan* p0, const mm_rate_plan* p1);
int mm_rate_plan_is_specialized_rf(const mm_rate_plan* p0, const mm_rate_plan* p1);
long mm_rate_plan_spec_source_rf(const mm_rate_plan* p0, const mm_rate_plan* p1, char* buf, long size);

/*
 * The same generated source from the tables of mm_rate_plan_create, without
 * a device (no allocation; for inspection and for building the kernels
 * ahead of time).  Same return convention as mm_rate_plan_spec_source.
 */
long mm_rate_spec_source_tables(int n_ground, int n_excited, int n_beams, const double* energies,
                                const double* zeeman_z_diag, const double* d_squared, const double* beam_dirs,
                                const double* beam_freqs, const int* beam_pols, const double* beam_s0s,
                                const int* beam_comp, const double* comp_split_sign, int n_comp, double Gamma,
                                double k_wave, double Gamma_eff, double hbar, int kind, char* buf, long size);

/*
 * Compiles a generated source with hipRTC for `arch` (e.g. "gfx950"), with
 * the library's device headers and flags; needs no device.  Copies the code
 * object (ELF) into `code` (truncated to size; may be NULL) and returns its
 * size, or a negative MM_* code with the compiler log in mm_last_error().
 */
long mm_rtc_compile(const char* source, const char* arch, char* code, long size);

/*
 * Rate-equation force map over the (v, z) grid.
 * Replaces solve_rate_equations_batch + rate_eq_kernel
 * (molmot_cuda/rate_equations_cuda.py:448-544, :51-277) and the per-point
 * loops of force_vs_z / force_vs_v (molmot/mot/force_scan.py:18-81).
 *   v[n_v], z[n_z]                      device
 *   force[n_v*n_z]                      device, index iv*n_z + iz (N)
 *   pop[n_v*n_z*n_ground] or NULL       device, ground populations
 *   scatter[n_v*n_z] or NULL            device, scattering rate (1/s)
 */
int mm_rate_eq_map(const mm_rate_plan* plan, const double* v, int n_v, const double* z, int n_z,
                   double B_gradient, double* force, double* pop, double* scatter, void* stream);

/*
 * Rate-equation force at n scattered points (v[i], z[i]) (all device).
 * Replaces DCMOTSimulator.force / solve_rate_equations_auto
 * (molmot/mot/simulator.py:132-154, molmot/obe/rate_equations.py:205-218)
 * for batches; pop / scatter may be NULL.
 */
int mm_rate_eq_points(const mm_rate_plan* plan, long n, const double* v, const double* z,
                      double B_gradient, double* force, double* pop, double* scatter, void* stream);

/*
 * Force and its gradient at n scattered points (all device; no reference
 * counterpart: the reference differentiates only by two-point finite
 * differences, molmot/mot/force_scan.py:84-128, and its optimiser scans a
 * grid, molmot/mot/force_scan.py:191-301).
 *   force[n] (may be NULL)
 *   jac[n][9]: dF/d(detuning_c) for frequency components c = 0..3 (N/Hz;
 *              the detuning of component c shifts every beam with that
 *              component index), dF/d(split) (N/Hz; beam frequency moves by
 *              comp_split_sign[c] * split), dF/d(s0 scale) (N; all s0
 *              multiplied by a common factor, at 1), dF/dB' (N m/T),
 *              dF/dz (N/m), dF/dv (N s/m).
 * Points where a ground state is dark (the force falls back to the dense
 * elimination) get NaN gradients.  Requires at most 4 components.
 */
int mm_rate_eq_grad_points(const mm_rate_plan* plan, long n, const double* v, const double* z,
                           double B_gradient, double* force, double* jac, void* stream);

/*
 * n independent 1D trajectories, symplectic Euler
 *   v_j = v_{j-1} + F(v_{j-1}, z_{j-1})/m dt,  z_j = z_{j-1} + v_j dt
 * for j = 1 .. n_steps-1, stopping a particle at the first |z_j| > z_escape
 * (n_actual = j; saved samples after it repeat the escape state).
 * Replaces simulate_trajectories_cuda / traj_kernel
 * (molmot_cuda/trajectories_cuda.py:305-426, :44-279) and
 * _trajectory_1d_jit (molmot/propagation/trajectories_jit.py:18-58).
 *   z0, v0 [n]                          device
 *   z_traj, v_traj [n*n_steps] or NULL  device (both or neither)
 *   z_final, v_final [n], n_actual [n]  device
 * Recoil heating (no reference counterpart): when kick_scale > 0 each step
 * adds kick_scale * sqrt(R dt) * xi to v, xi ~ N(0,1) drawn from
 * Philox4x32-10 with key (seed_lo, seed_hi) and counter (step / 2, particle
 * low 32 bits, stream_id, particle high 32 bits): one block gives a
 * Box-Muller pair, its cosine normal for odd steps and its sine normal for
 * the following even step; R is the scattering rate at the step's force
 * evaluation.  kick_scale = v_recoil * sqrt(1 + 1/3) models
 * absorption along +-z plus isotropic emission projected on z.
 */
int mm_traj_euler_1d(const mm_rate_plan* plan, long n, const double* z0, const double* v0,
                     double B_gradient, double mass, double dt, int n_steps, double z_escape,
                     double* z_traj, double* v_traj, double* z_final, double* v_final,
                     int* n_actual, double kick_scale, unsigned int seed_lo,
                     unsigned int seed_hi, unsigned int stream_id, void* stream);

/*
 * Bytes of workspace mm_traj_euler_1d_ws needs to schedule n trajectories
 * of n_steps through its work queue (256 + 8 x groups x (chunks - 1), with
 * groups = ceil(n / 64) and at most 256 chunks of at least 128 steps), or 0
 * if the queue does not apply (n_steps <= 257).
 */
long mm_traj_euler_workspace_bytes(long n, int n_steps);

/*
 * mm_traj_euler_1d with a choice of schedule; the results are bit-identical
 * (same arithmetic per step, same Philox counters).  No reference
 * counterpart: the reference launches one thread per particle
 * (molmot_cuda/trajectories_cuda.py:305-426), which is schedule 1 here.
 *   schedule 1 (grid): one thread per particle, as mm_traj_euler_1d.
 *   schedule 2 (queue): a persistent grid of wavefronts takes (group of 64
 *     particles, chunk of steps) units from a queue in `workspace`, so that
 *     every SIMD carries the same share of the ensemble -- one thread per
 *     particle leaves 1e5 particles as 1,563 wavefronts on 1,024 SIMDs.
 *     Needs z_traj == NULL and workspace_bytes >= mm_traj_euler_workspace_bytes.
 *   schedule 0 (auto): the queue when it applies and one thread per
 *     particle would leave its last round of wavefronts under 90% full
 *     (over the SIMDs, or over all resident slots when the groups do not
 *     fit at once); else the grid.  The queue gains nothing where the
 *     grid is balanced and gains 17% at 1e5 particles.
 * workspace: device memory, 256-byte aligned, cleared by a kernel on
 * `stream` at each launch (no memset: the launch can be captured into a
 * HIP graph); launches that share a workspace must not overlap.  A queue
 * wait that exceeds max(10 s, 10 ms x steps per chunk) (never in a correct
 * run; mm_traj_queue_set_wait overrides the bound) ends every wavefront,
 * sets n_actual to -1 for all particles and leaves a non-zero 32-bit word
 * at byte offset 8 of the workspace.
 */
/*
 * The schedule mm_traj_euler_1d_ws runs for schedule 0 (auto) on the
 * current device: 1 (grid) or 2 (queue), given a large enough workspace;
 * negative on error.  No reference counterpart (a query for callers and
 * benchmarks that report which kernel ran).
 */
int mm_traj_euler_schedule(const mm_rate_plan* plan, long n, int n_steps, int save_full, double kick_scale);

int mm_traj_euler_1d_ws(const mm_rate_plan* plan, long n, const double* z0, const double* v0,
                        double B_gradient, double mass, double dt, int n_steps, double z_escape,
                        double* z_traj, double* v_traj, double* z_final, double* v_final, int* n_actual,
                        double kick_scale, unsigned int seed_lo, unsigned int seed_hi, unsigned int stream_id,
                        int schedule, void* workspace, long workspace_bytes, void* stream);

/*
 * Batched dense complex FP64 solves A_s X_s = B_s (s < batch), in place:
 * A (batch, n, n) and B (batch, n, nrhs) complex128 interleaved, row-major,
 * device memory; A is overwritten by its LU factors (unit-lower L below the
 * diagonal), B by X.  Gaussian elimination with partial pivoting (LAPACK's
 * pivot choice), one workgroup per system.  Replaces cp.linalg.solve in the
 * batched steady state (molmot_cuda/lindblad_cuda.py:236-242) and the dense
 * solve of the Floquet system (molmot/obe/floquet.py:288-293,
 * molmot_cuda/floquet_cuda.py:295).  Optional outputs (NULL = not wanted):
 * info[s] = k + 1 for the first exactly-zero pivot (0 if none), umin[s] /
 * umax[s] = min / max of |Re U_kk| + |Im U_kk|.  n <= 2048.
 */
int mm_zgesv_batched(int n, int nrhs, long batch, double* A, double* B, int* info, double* umin, double* umax,
                     void* stream);

/*
 * Ensemble statistics of the 1D trajectories at sample steps, computed on
 * the device without storing trajectories (replaces the host loops of
 * sigma_vs_time / temperature_vs_time / density_vs_time,
 * molmot/mot/analysis.py:128-288, over saved (t, z, v) lists).  The
 * particles follow mm_traj_euler_1d exactly (same steps, escape rule and
 * Philox kicks; escaped particles keep their escape state, as the
 * reference's escape fill does).  sample_steps: device int array of n_samp
 * non-decreasing state indices in [0, n_steps); at each, the particles
 * with |z| < r_alive contribute.  stats (device, n_samp x 5): alive count,
 * mean z, sum of (z - mean z)^2, mean v, sum of (v - mean v)^2, merged
 * pairwise in a fixed order (deterministic).  workspace: device memory of
 * mm_traj_stats_workspace_bytes(n, n_samp) bytes.
 */
long mm_traj_stats_workspace_bytes(long n, int n_samp);
int mm_traj_stats_1d(const mm_rate_plan* plan, long n, const double* z0, const double* v0, double B_gradient,
                     double mass, double dt, int n_steps, double z_escape, double r_alive, const int* sample_steps,
                     int n_samp, double* stats, double kick_scale, unsigned int seed_lo, unsigned int seed_hi,
                     unsigned int stream_id, void* workspace, long workspace_bytes, void* stream);

/*
 * mm_traj_euler_1d_ws (finals only, no saved trajectory) without a caller
 * workspace.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It allocates nothing: the work queue
 * lives in a file-scope device array of mm_traj_euler_noalloc_max_bytes()
 * bytes (4 MiB: up to 256 + 8 x groups x (chunks - 1) bytes, e.g. 131,072
 * particles x 32,768 steps), held by every GPU the library runs on until the
 * process ends.  Sizes whose queue would not fit return MM_EINVAL unless
 * schedule is 1 (one thread per particle, no queue).  Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (other streams,
 * other host threads) race; the array keeps the last call's contents, and
 * each launch clears the part it uses with a kernel first.  Results are
 * bit-identical to mm_traj_euler_1d_ws with the same schedule.
 */
long mm_traj_euler_noalloc_max_bytes(void);
int mm_traj_euler_1d_noalloc(const mm_rate_plan* plan, long n, const double* z0, const double* v0,
                             double B_gradient, double mass, double dt, int n_steps, double z_escape,
                             double* z_final, double* v_final, int* n_actual, double kick_scale,
                             unsigned int seed_lo, unsigned int seed_hi, unsigned int stream_id, int schedule,
                             void* stream);

/*
 * Bound of a work-queue wait, in seconds, for every later queue launch of
 * the process (mm_traj_euler_1d_ws, mm_traj_tangent_1d_ws, mm_traj_vsq_1d_ws,
 * mm_traj_axis3d_ws); seconds < 0 restores the default max(10 s, 10 ms x
 * steps per chunk).  No reference counterpart (the reference has no queue):
 * it lets a caller that time-shares the GPU allow longer waits, and lets a
 * test drive the give-up path with a bound of 0 while every kernel still
 * exits normally.  MM_EINVAL for a bound above one day or not a number.
 */
int mm_traj_queue_set_wait(double seconds);

/*
 * RF MOT trajectories: as mm_traj_euler_1d with the force and scattering
 * rate of each step the mean of plan0 at +B_gradient and plan1 at
 * -B_gradient (the two switching phases, molmot/mot/simulator.py:49-90,
 * integrated as molmot/propagation/trajectories.py:19-75 does for any
 * simulator; the reference's GPU path falls back to the CPU for RF,
 * molmot_cuda/force_scan_cuda.py:578-585).  Plans of equal size, beams along +-z.
 */
int mm_traj_euler_rf_1d(const mm_rate_plan* plan0, const mm_rate_plan* plan1, long n, const double* z0,
                        const double* v0, double B_gradient, double mass, double dt, int n_steps, double z_escape,
                        double* z_traj, double* v_traj, double* z_final, double* v_final, int* n_actual,
                        double kick_scale, unsigned int seed_lo, unsigned int seed_hi, unsigned int stream_id,
                        void* stream);

/*
 * 3D per-axis trajectories (replaces simulate_trajectories_3d,
 * molmot/propagation/trajectories.py:78-156): each step evaluates the 1D
 * force at (v_z, z), (v_x, -x/2), (v_y, -y/2) -- the quadrupole field
 * B'(-x/2, -y/2, z) -- then v += F/m dt, r += v dt, and stops a particle at
 * the first |r| > r_escape (n_actual = j; saved samples repeat the escape
 * state).  Four lanes per particle (one per axis, one idle).
 *   r0, v0, r_final, v_final [n][3]; r_traj, v_traj [n][n_steps][3] or NULL
 */
int mm_traj_axis3d(const mm_rate_plan* plan, long n, const double* r0, const double* v0,
                   double B_gradient, double mass, double dt, int n_steps, double r_escape,
                   double* r_traj, double* v_traj, double* r_final, double* v_final,
                   int* n_actual, void* stream);

/*
 * Workspace bytes of mm_traj_axis3d_ws's queue (groups of 16 particles, one
 * per wavefront: 256 + 8 x ceil(n/16) x (chunks - 1)), or 0 if the queue
 * does not apply (n_steps <= 257).
 */
long mm_traj_axis3d_workspace_bytes(long n, int n_steps);

/*
 * mm_traj_axis3d with a choice of schedule, as mm_traj_euler_1d_ws: 1 (one
 * lane quad per particle), 2 (work queue of (16-particle group, chunk of
 * steps) units in `workspace`; needs r_traj == NULL), 0 (auto, the same
 * rule).  Bit-identical results.  A timed-out queue wait (never in a correct
 * run) sets every n_actual to -1.  No reference counterpart (the reference
 * integrates the per-axis model on the CPU).
 */
int mm_traj_axis3d_ws(const mm_rate_plan* plan, long n, const double* r0, const double* v0, double B_gradient,
                      double mass, double dt, int n_steps, double r_escape, double* r_traj, double* v_traj,
                      double* r_final, double* v_final, int* n_actual, int schedule, void* workspace,
                      long workspace_bytes, void* stream);

/*
 * Monte-Carlo temperature statistics (replaces the particle loop of
 * monte_carlo_temperature, molmot/propagation/trajectories.py:159-229, and
 * _monte_carlo_temperature_jit, trajectories_jit.py:61-109): Euler steps
 * j = 1 .. n_steps-1; a particle stops after the first step with
 * |z| > z_escape; at every other step with j >= i_eq, v^2 is summed.
 *   vsq_sum[n] (double), n_samples[n] (int64), per particle (device)
 */
int mm_traj_vsq_1d(const mm_rate_plan* plan, long n, const double* z0, const double* v0,
                   double B_gradient, double mass, double dt, int n_steps, int i_eq,
                   double z_escape, double* vsq_sum, long long* n_samples, void* stream);

/*
 * Workspace bytes of mm_traj_vsq_1d_ws's queue (the queue of
 * mm_traj_euler_workspace_bytes rounded up to 256, then z and v of every
 * particle: 16 n), or 0 if the queue does not apply (n_steps <= 257).
 */
long mm_traj_vsq_workspace_bytes(long n, int n_steps);

/*
 * mm_traj_vsq_1d with a choice of schedule, as mm_traj_euler_1d_ws: 1 (one
 * thread per particle), 2 (work queue in `workspace`), 0 (auto, the same
 * rule).  Bit-identical results.  A timed-out queue wait (never in a correct
 * run) sets every n_samples to -1.  No reference counterpart (the
 * reference loops over particles on the CPU).
 */
int mm_traj_vsq_1d_ws(const mm_rate_plan* plan, long n, const double* z0, const double* v0, double B_gradient,
                      double mass, double dt, int n_steps, int i_eq, double z_escape, double* vsq_sum,
                      long long* n_samples, int schedule, void* workspace, long workspace_bytes, void* stream);

/*
 * RF MOT versions of mm_traj_axis3d and mm_traj_vsq_1d: every force
 * evaluation is the mean of plan0 at +B_gradient and plan1 at -B_gradient
 * (the switching phases, molmot/mot/simulator.py:49-90); the reference runs
 * these through the CPU integrators for any simulator
 * (molmot/propagation/trajectories.py:78-229).  Plans of equal size, beams
 * along +-z.
 */
int mm_traj_axis3d_rf(const mm_rate_plan* plan0, const mm_rate_plan* plan1, long n, const double* r0,
                      const double* v0, double B_gradient, double mass, double dt, int n_steps, double r_escape,
                      double* r_traj, double* v_traj, double* r_final, double* v_final, int* n_actual, void* stream);
int mm_traj_vsq_rf_1d(const mm_rate_plan* plan0, const mm_rate_plan* plan1, long n, const double* z0,
                      const double* v0, double B_gradient, double mass, double dt, int n_steps, int i_eq,
                      double z_escape, double* vsq_sum, long long* n_samples, void* stream);

/*
 * Standing-wave Lindblad system class (device tables).  Replaces the per-call
 * Hamiltonian / Liouvillian construction of obe_force_subdoppler
 * (molmot/obe/obe_subdoppler.py:19-141) and build_liouvillian_batch
 * (molmot_cuda/lindblad_cuda.py:124-196).  All in natural units (Gamma = 1):
 *   Hd[n_g+n_e]                 diagonal of H (multi-rotating frame)
 *   A_re/A_im, B_re/B_im        [n_e][n_g] forward / backward couplings
 *   gam[n_e]                    decay rate of each excited state
 *   rep[n_g][n_e]               repopulation rate e -> g
 *   force_unit = 2 hbar k Gamma (N), rate_unit = Gamma_eff (1/s)
 * Host pointers; allocates the device copy.  n_g <= 36, n_e <= 4.
 */
typedef struct mm_lindblad_class mm_lindblad_class;
int mm_lindblad_class_create(mm_lindblad_class** cls, int n_ground, int n_excited, const double* Hd,
                             const double* A_re, const double* A_im, const double* B_re,
                             const double* B_im, const double* gam, const double* rep,
                             double force_unit, double rate_unit);
int mm_lindblad_class_destroy(mm_lindblad_class* cls);

/*
 * Batched RK4 propagation of n_inst density matrices of one class, one
 * wavefront per instance, with the standing-wave phase phi0[i] + kv[i] t
 * (kv = k v / Gamma).  Starts from equal ground populations, takes n_steps
 * steps of dt (1/Gamma) and averages the force and excited population over
 * the step ends after step avg_start.  No reference counterpart: the
 * reference solves the steady state at fixed phase (obe_subdoppler.py:144-183)
 * and never uses v.
 *   precision 0: FP64, VALU contraction using the block structure of H
 *             1: FP32, same contraction
 *             2: FP32, H rho on MFMA (12+4 only; 10 v_mfma_f32_16x16x4f32
 *                per right-hand side)
 *   F[n_inst] (N), R[n_inst] (1/s, may be NULL),
 *   rho_final[n_inst][N][N][2] (re, im; FP64 for precision 0, else FP32; may be NULL)
 */
int mm_lindblad_rk4(const mm_lindblad_class* cls, long n_inst, const double* kv, const double* phi0,
                    double dt, int n_steps, int avg_start, int precision, double* F, double* R,
                    void* rho_final, void* stream);

/*
 * Lanes per instance of the VALU back end (precision 0 or 1): 64, 128 or 256
 * (one, two or four wavefronts share an instance, each holding fewer of its
 * N^2 elements), or 0 for the default (64 up to 16 levels, else 256).  A
 * tuning knob with no reference counterpart; molmot_amd.autotune picks it by
 * timing.  get_lanes returns the value in effect (negative on error).
 */
int mm_lindblad_class_set_lanes(mm_lindblad_class* cls, int precision, int lanes);
int mm_lindblad_class_get_lanes(const mm_lindblad_class* cls, int precision);

/*
 * Uneven batch in one launch: n16 instances of a 12+4 class and nbig of a
 * 36+4 class, each with its own kv, phi0 and outputs as in mm_lindblad_rk4
 * (one shared dt, n_steps, avg_start).  precision 0: FP64 VALU for both;
 * 2: FP32, the 16-level part on MFMA and the 40-level part on VALU.  The
 * 40-level instances take four wavefronts each, the 16-level ones one
 * wavefront each, four per block.  No reference counterpart (the reference
 * runs one class per solve, molmot_cuda/lindblad_cuda.py:448-538).
 */
int mm_lindblad_rk4_uneven(const mm_lindblad_class* c16, long n16, const double* kv16, const double* phi16,
                           double* F16, double* R16, void* rho16, const mm_lindblad_class* cbig, long nbig,
                           const double* kvbig, const double* phibig, double* Fbig, double* Rbig, void* rhobig,
                           double dt, int n_steps, int avg_start, int precision, void* stream);

/*
 * Augmented Liouvillians for the steady-state solve: for each instance
 * (standing-wave phase phi0[i]) the (N^2 x N^2) superoperator of
 * build_liouvillian_fast (molmot/obe/obe_subdoppler.py:107-141; column-major
 * vec) in natural units; with trace_row != 0 its last row is replaced by the
 * trace constraint as steady_state_batch does
 * (molmot_cuda/lindblad_cuda.py:224-232).
 *   out[n_inst][N^2][N^2][2] (re, im), device; n_inst <= 65535
 */
int mm_lindblad_liouvillian(const mm_lindblad_class* cls, long n_inst, const double* phi0, int trace_row,
                            double* out, void* stream);

/*
 * Reverse-mode gradient through the 1D trajectory integrator without recoil
 * kicks (the integrator of mm_traj_euler_1d; molmot/propagation/trajectories.py:19-75).
 * No counterpart in the reference, which has no derivatives; it serves the
 * capture-velocity gradient in place of the grid search of
 * molmot/mot/force_scan.py:131-188.
 * Particle i runs min(n_up[i], n_steps - 1) updates (all n_steps - 1 if
 * n_up is NULL or n_up[i] < 0), stopping early when |z| > z_escape; with
 * L_i = cot_z[i] z_N + cot_v[i] v_N at its last state N (cot_* NULL = 0) it
 * returns z_final, v_final, n_done (= N), g_theta[n][7] = dL_i/d[det0..3,
 * split, s0 scale, B'] and g_z0, g_v0 = dL_i/dz_0, dL_i/dv_0.
 * Checkpointed recomputation every K states; work holds
 * mm_traj_vjp_workspace_doubles(n, n_steps, K) doubles (device).
 */
long mm_traj_vjp_workspace_doubles(long n, int n_steps, int K);
int mm_traj_vjp_1d(const mm_rate_plan* plan, long n, const double* z0, const double* v0, double B_gradient,
                   double mass, double dt, int n_steps, const int* n_up, double z_escape, int K,
                   const double* cot_z, const double* cot_v, double* z_final, double* v_final, int* n_done,
                   double* g_theta, double* g_z0, double* g_v0, double* work, long work_doubles, void* stream);

/*
 * Forward sensitivities through the same integrator (no counterpart in the
 * reference, which has no derivatives; it serves the capture-velocity and
 * trajectory-loss gradients that replace the grid searches of
 * molmot/mot/force_scan.py:131-188 and :191-301).  One pass; per particle
 * sens[i][0][j] = d z_N / d p_j and sens[i][1][j] = d v_N / d p_j for
 * p = (det0..3, split, s0 scale, B', z_0, v_0), N and the stopping rule as
 * in mm_traj_vjp_1d.  Allocates nothing; no workspace.
 */
int mm_traj_tangent_1d(const mm_rate_plan* plan, long n, const double* z0, const double* v0, double B_gradient,
                       double mass, double dt, int n_steps, const int* n_up, double z_escape, double* z_final,
                       double* v_final, int* n_done, double* sens, void* stream);

/*
 * mm_traj_tangent_1d with a choice of schedule, as mm_traj_euler_1d_ws:
 * 1 (grid, one thread per particle), 2 (queue: a persistent grid takes
 * (64-particle group, chunk of updates) units from a queue in `workspace`,
 * sized by mm_traj_euler_workspace_bytes(n, n_steps)), 0 (auto, the same
 * rule).  Bit-identical results.  A timed-out queue wait (never in a correct
 * run) sets n_done to -1 for every particle.  No reference counterpart
 * (the reference has no trajectory gradient).
 */
int mm_traj_tangent_1d_ws(const mm_rate_plan* plan, long n, const double* z0, const double* v0, double B_gradient,
                          double mass, double dt, int n_steps, const int* n_up, double z_escape, double* z_final,
                          double* v_final, int* n_done, double* sens, int schedule, void* workspace,
                          long workspace_bytes, void* stream);

/* ---------------------------------------------------------------------------
 * Monte Carlo wavefunction (quantum jump) ensemble.
 * Replaces SSESolverCUDA.run_ensemble / run_ensemble_cuda
 * (molmot_cuda/stochastic_cuda.py:938-1173) and the per-trajectory loop of
 * SSESolver.run (molmot/obe/stochastic.py:661-804); the tables are those of
 * SSEProblem (stochastic.py:321-440) and _prepare_problem_buffers
 * (stochastic_cuda.py:786-858), in natural units (t Gamma, r k, v k/Gamma).
 * ------------------------------------------------------------------------- */
typedef struct mm_mcwf_problem mm_mcwf_problem;

typedef struct {
    double dt;              /* step (natural units) */
    double r_escape;        /* stop when |r| > r_escape (0 = never) */
    int n_steps;
    int save_every;         /* save at steps 0, save_every, ... (<= n_save points) */
    int n_save;
    int renorm_interval;    /* 0 = never (stochastic_cuda.py:729-748) */
    int avg_start;          /* force_sum covers steps >= avg_start */
    int precision;          /* 0: FP64 amplitudes, 1: FP32 amplitudes (norm, r, v FP64) */
    int jump_cap;           /* jumps logged per trajectory (0 = none) */
    int jump_rule;          /* 0: jump when 1-|psi|^2 > -ln U (reference rule);
                               1: when |psi|^2 < U (exact waiting time) */
    unsigned long long seed;
} mm_mcwf_config;

/* Device outputs; optional ones may be NULL.  Shapes: r, v, force
 * [n][n_save][3]; pop [n][n_save][N]; force_sum, r_final, v_final [n][3];
 * psi_final [n][N][2]; jump_t, jump_ch [n][jump_cap] (channel = (e n_g + g) 3 + q). */
typedef struct {
    double *r, *v, *pop, *force;
    double *force_sum, *psi_final, *r_final, *v_final, *jump_t;
    int *n_scatters, *escaped, *n_saved, *n_avg, *jump_ch;
} mm_mcwf_outputs;

/*
 * Problem tables (host arrays, copied): wbase[N] energies minus the ground /
 * excited reference; zeeman_re/im[3][n_g][n_g] Hermitian ground blocks
 * (x, y, z); zdiag_e[n_e]; d[n_g][n_e][3] = |tdm|; br[n_e][n_g][3];
 * amp_re/im[6][n_f][3] = sqrt(s_f)/(2 sqrt 2) conj(eps_{b,f,q}) for beams
 * +x +y +z -x -y -z; w_lasers[n_f] minus (excited - ground reference).
 * Sizes (n_g, n_e) in {(12,4), (1,1), (2,2)}; n_f <= 8.
 */
int mm_mcwf_problem_create(mm_mcwf_problem** out, int n_ground, int n_excited, int n_freq, const double* wbase,
                           const double* zeeman_re, const double* zeeman_im, const double* zdiag_e,
                           const double* d, const double* br, const double* amp_re, const double* amp_im,
                           const double* w_lasers, double B_gradient, double r_nat, double beam_radius_k,
                           double mass_nat, double v_recoil, int add_kick, const double* diffusion);
int mm_mcwf_problem_destroy(mm_mcwf_problem* p);

/* Laser sums at the stage times t_tau (device, [n_tau]): S[n_tau][6][3]
 * complex (double2 for precision 0, float2 for 1).  For a run of n_steps the
 * table holds t_0, t_0 + dt/2, t_1, t_1 + dt/2, ..., t_n, t_n + dt/2 with
 * t_{k+1} = t_k + dt accumulated in FP64. */
int mm_mcwf_field_table(const mm_mcwf_problem* p, long n_tau, const double* t_tau, int precision, void* S,
                        void* stream);

/* n trajectories: r0, v0 [n][3], psi0 [n][N][2] (NULL = first ground state). */
int mm_mcwf_run(const mm_mcwf_problem* p, long n, const double* r0, const double* v0, const double* psi0,
                const void* field_table, long n_tau, const mm_mcwf_config* cfg, const mm_mcwf_outputs* out,
                void* stream);

/* ---------------------------------------------------------------------------
 * 3D rate-equation force with the full ground Zeeman Hamiltonian.
 * Replaces rate_eq_force_3d (molmot/mot/simulator_3d.py:168-365),
 * rate_eq_force_3d_batch (molmot_cuda/simulator_3d_cuda.py:666-754) and the
 * leapfrog of simulate_trajectory_3d / simulate_ensemble_3d(_cuda)
 * (simulator_3d.py:439-532, simulator_3d_cuda.py:818-973).
 * ------------------------------------------------------------------------- */
typedef struct mm_force3d_plan mm_force3d_plan;

/* Host tables (copied): energies[N] Hz; zeeman_re/im[3][n_g][n_g] ground
 * blocks (x, y, z); zdiag_e[n_e] Re diag of the excited z block;
 * dabs[n_g][3 n_e] = |tdm[g, n_g+e, q]| at column 3e+q; beams: k[nb][3] unit
 * directions, freq[nb] Hz, q[nb] in {0,1,2} (lab frame), s0[nb], w[nb] 1/e^2
 * radius in m (<= 0: no profile).  (n_g, n_e) in {(12,4), (1,1), (2,2)}, nb <= 48. */
int mm_force3d_plan_create(mm_force3d_plan** out, int n_ground, int n_excited, int n_beams, const double* energies,
                           const double* zeeman_re, const double* zeeman_im, const double* zdiag_e,
                           const double* dabs, const double* beam_k, const double* beam_freq, const int* beam_q,
                           const double* beam_s0, const double* beam_w, double Gamma, double k, double Geff,
                           double hbar, double mass);
int mm_force3d_plan_destroy(mm_force3d_plan* plan);

/* Spin-rotation eigenbasis for ground Zeeman blocks that are s times spin
 * operators of a sum of spin-1/2 representations (checked by the caller):
 * lam0[n_g] ascending eigenvalues of the Hermitised z block, W its
 * eigenvectors; sd[n_g][3 n_e][6] = (W^H |d|, W^H (2/s) Zx |d|,
 * W^H (2/s) Zy |d|) as (re, im) pairs; inv_tot[n_e] = 1 / sum_{g,q} |d|^2 (0
 * where that sum is <= 1e-30).  The force kernels then take the eigenvectors
 * of H_Z = B.Z as R(B) W (R the spin rotation from z to B / |B|), one lane
 * per point, instead of diagonalising H_Z at every point as
 * rate_eq_force_3d does (molmot/mot/simulator_3d.py:230-250).  Where the
 * Zeeman spectrum is degenerate this is the basis convention; lam0 = NULL
 * returns to the per-point diagonalisation. */
int mm_force3d_plan_set_spin_basis(mm_force3d_plan* plan, const double* lam0, const double* sd,
                                   const double* inv_tot);

/* The leapfrog of mm_force3d_leapfrog in segments of seg_steps (0: 256)
 * steps, for plans with the spin-rotation basis (others run
 * mm_force3d_leapfrog).  After each segment the particles still inside
 * r_escape are compacted into a list (order kept) and the next segment runs
 * on them alone, with 1, 4 or 16 lanes per particle (the most that keep
 * them at <= 2 wavefronts per SIMD), so that the last particles inside do
 * not leave wavefronts of idle lanes running.  r_final, v_final, n_actual
 * are required (they carry the state between segments).  Workspace:
 * mm_force3d_leapfrog_workspace_bytes(n) bytes, 256-byte aligned.  The
 * stream is synchronised once per segment (the alive count sizes the next
 * launch).  Results equal mm_force3d_leapfrog's up to the order of the
 * per-ground-state sums in segments with more than one lane per particle. */
long mm_force3d_leapfrog_workspace_bytes(long n);
int mm_force3d_leapfrog_ws(const mm_force3d_plan* plan, long n, const double* r0, const double* v0,
                           double B_gradient, double dt, int n_steps, double r_escape, double* r_traj,
                           double* v_traj, double* r_final, double* v_final, int* n_actual, int seg_steps,
                           void* ws, long ws_bytes, void* stream);

/* F[n][3] (N), R[n] (rad/s, may be NULL), pop[n][n_g] (may be NULL) at
 * positions r[n][3] (m) and velocities v[n][3] (m/s); device pointers. */
int mm_force3d_points(const mm_force3d_plan* plan, long n, const double* r, const double* v, double B_gradient,
                      double* F, double* R, double* pop, void* stream);

/* Leapfrog ensemble of n particles, n_steps stored steps (index 0 = initial
 * state).  r_traj/v_traj [n][n_steps][3] may both be NULL; r_final, v_final
 * [n][3] (position and half-step velocity), n_actual[n] = first escaped step
 * (n_steps if none); any of the three may be NULL. */
int mm_force3d_leapfrog(const mm_force3d_plan* plan, long n, const double* r0, const double* v0, double B_gradient,
                        double dt, int n_steps, double r_escape, double* r_traj, double* v_traj, double* r_final,
                        double* v_final, int* n_actual, void* stream);

/*
 * Philox4x32-10 block function (Salmon et al., SC'11) on n counters, for
 * known-answer tests of the generator the trajectory kernels use.
 *   ctr[4*n] (device, uint32), out[4*n] (device, uint32)
 */
int mm_philox4x32_10(long n, const unsigned int* ctr, unsigned int key0, unsigned int key1,
                     unsigned int* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MOLMOT_HIP_H */
