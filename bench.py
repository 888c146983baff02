#!/usr/bin/env python3
"""Benchmark: 1D rate-equation trajectory ensemble on one MI355X.

Workload (one "step"): 100,000 SrOH molecules (12+4 levels, bundled tables)
x 30,000 symplectic-Euler steps of dt = 10 us in the four-frequency red MOT
(detunings [-10.3, -5.9, -9.2, -29.9] MHz, B' = 21.4 G/cm, 100 mW over six
beams of 5 mm waist), FP64 state and force, photon-recoil kicks from
Philox4x32-10.  Initial conditions: z0 ~ N(0, 1 mm), v0 = Maxwell-Boltzmann
(5 mK) + U[-15, 15] m/s, seed 42, resident in HBM before timing.

value = N_particles * n_steps / (wall time per step)  [particle-steps/s].
Also reported: the 1024 x 1024 FP64 force map (force-pts/s), the dominant
kernel's achieved FP64 rate against the MI355X peak (HIP events on the
launch stream), and the oracle (C FP64 restatement of the reference's
Numba loop) timed on this host's cores on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "particle-steps/s & force-pts/s (fwd+adjoint, ≤1e-8 rel err vs FP64 CPU)"
FLOP_PER_EVAL = 4268          # one 12+4, 8-beam rate-equation evaluation (FMA = 2)
FLOP_PER_PARTICLE_STEP = 4272  # + the Euler update
PEAK_FP64_TFLOPS = 78.6        # MI355X FP64 (vector and matrix peaks are equal)


def _host():
    """(threads to use, nproc, CPU model): the cores this process may run on,
    capped by OMP_NUM_THREADS when it is set (a GPU host shares its cores
    between the GPUs' jobs and says so there)."""
    try:
        n_aff = len(os.sched_getaffinity(0))
    except AttributeError:
        n_aff = os.cpu_count() or 1
    cores = n_aff
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        cores = min(cores, int(omp))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return max(1, cores), os.cpu_count() or 1, n_aff, model


def _progress(msg):
    """One line per finished section on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_sample_size(cores):
    return 64 * cores


def cpu_baseline(mol, beams, z0, v0, B, dt, n_steps):
    """Oracle (reference algorithm restated in C, OpenMP) on host cores.
    Sample: the first 64 * cores particles of the same ensemble, all 30,000
    steps, no recoil kicks; reported in the same unit (nominal
    particle-steps / s).  Returns (baseline dict, oracle finals)."""
    import oracle
    from molmot_amd.constants import hbar
    from molmot_amd.obe.fields import beam_table
    cores, nproc, n_aff, model = _host()
    isa = oracle.use_avx512_build()
    n = cpu_sample_size(cores)
    bt = beam_table(beams)
    oracle.lib()
    t0 = time.perf_counter()
    zf, vf, na, _, _ = oracle.traj_euler(mol, bt, z0[:n], v0[:n], dt, n_steps, B, hbar=hbar, n_threads=cores)
    el = time.perf_counter() - t0
    return ({"value": n * n_steps / el, "unit": "particle-steps/s", "cores": cores, "kind": "port",
             "sample": f"{n} particles x {n_steps} steps of the same ensemble (no recoil kicks), "
                       f"{el:.1f} s on {cores} threads",
             "host": {"nproc": nproc, "cores_in_affinity": n_aff, "cpu_model": model,
                      "threads_used": cores, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
                      "oracle_build": f"gcc -O3 -march={isa} -fopenmp"}},
            (zf, vf, na))


def ensemble_parity(plan, z0, v0, B, dt, n_steps, ref, stream):
    """The headline ensemble's error against the FP64 oracle: the GPU
    kick-free forward of the particles the CPU leg integrated, compared
    per particle (final z and v, escape step)."""
    from molmot_amd.hip import traj_euler_device
    import torch
    zo, vo, nao = ref
    n = zo.size
    zf, vf, na, _, _ = traj_euler_device(plan, z0[:n], v0[:n], B, dt, n_steps, 0.015, stream=stream)
    torch.cuda.synchronize()
    zf, vf, na = zf.cpu().numpy(), vf.cpu().numpy(), na.cpu().numpy()
    dz, dv = np.abs(zf - zo), np.abs(vf - vo)
    same = na == nao
    return {"sample": f"the CPU leg's {n} particles x {n_steps} steps, no kicks, FP64 oracle "
                      "(molmot/propagation/trajectories_jit.py:18-58 restated in C)",
            "max_abs_dz_m": float(dz.max()), "max_abs_dv_m_per_s": float(dv.max()),
            "max_rel_z_final": float(dz.max() / np.abs(zo).max()),
            "max_rel_v_final": float(dv.max() / np.abs(vo).max()),
            "p95_abs_dv_m_per_s": float(np.percentile(dv, 95)),
            "escape_step_identical_fraction": float(np.mean(same)),
            "escaped_fraction": float(np.mean(nao < n_steps))}


def map_parity(mol, beams, F, P, R, v_g, z_g, B, n_sample=4096, seed=3):
    """The 1024 x 1024 map's error on a seeded sample of grid points against
    the FP64 oracle (max-norm relative over the sample)."""
    import oracle
    from molmot_amd.constants import hbar
    from molmot_amd.obe.fields import beam_table
    rng = np.random.default_rng(seed)
    nz = z_g.size
    idx = rng.choice(v_g.size * nz, n_sample, replace=False)
    iv, iz = idx // nz, idx % nz
    Fo, Po, Ro = oracle.rate_eq_points(mol, beam_table(beams), v_g[iv], z_g[iz], B, hbar=hbar,
                                       n_threads=_host()[0])
    Fg, Pg, Rg = F[idx], P[idx], R[idx]
    return {"sample": f"{n_sample} seeded grid points vs the FP64 oracle (rate_equations_jit.py:18-145 in C)",
            "F_max_rel": float(np.max(np.abs(Fg - Fo)) / np.max(np.abs(Fo))),
            "populations_max_abs": float(np.max(np.abs(Pg - Po))),
            "R_max_rel": float(np.max(np.abs(Rg - Ro)) / np.max(np.abs(Ro))),
            "populations_sum_max_dev": float(np.max(np.abs(Pg.sum(1) - 1.0)))}


def adjoint_bench(mol, beams, z0, v0, B, dt, n_steps, stream, repeats=3):
    """Deterministic (kick-free) ensemble forward and its gradient on the same
    100,000 x 30,000 workload: loss = sum of final positions, gradient w.r.t.
    the 7 beam/field parameters and the initial state of every particle, by
    the checkpointed adjoint (K = 500: forward with checkpoints, then each
    segment recomputed with its force Jacobians and walked back) and by
    forward sensitivities (one pass carrying the 2 x 9 Jacobian of the state,
    which includes the forward result).  Both are timed against the plain
    forward (each the median of `repeats` runs); the two gradients are
    compared."""
    import torch
    from molmot_amd.adjoint import ForceProblem, trajectory_tangent, trajectory_vjp, vjp_from_sens
    from molmot_amd.hip import traj_euler_device
    prob = ForceProblem(mol, beams)
    theta = prob.theta0(B).numpy()
    plan = prob.plan(theta)
    n = z0.numel()
    ones = torch.ones(n, dtype=torch.float64, device=z0.device)
    # warm-up launches (module load, first-touch) of each timed kernel; the
    # forward and the gradients run the plan-specialised kernels of one plan,
    # built here once and passed to every timed call
    from molmot_amd.hip.plan import FORWARD, GRADIENT
    if plan.specializable(GRADIENT):
        plan.specialize(FORWARD | GRADIENT)
    elif plan.specializable(FORWARD):
        plan.specialize(FORWARD)
    traj_euler_device(plan, z0[:64], v0[:64], B, dt, 10, 0.015, stream=stream)
    trajectory_vjp(prob, theta, z0[:64], v0[:64], mol.mass, dt, 9, ones[:64], None, 0.015, None, 500, stream,
                   plan=plan)
    trajectory_tangent(prob, theta, z0[:64], v0[:64], mol.mass, dt, 9, 0.015, None, stream, plan=plan)
    torch.cuda.synchronize()

    def timed(fn):
        # median of three: single launches of this length vary by up to ~20%
        # with the device clock
        ts, out = [], None
        for _ in range(repeats):
            t0 = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        return sorted(ts)[len(ts) // 2], out

    t_fwd, _ = timed(lambda: traj_euler_device(plan, z0, v0, B, dt, n_steps, 0.015, stream=stream))
    t_fwd_grid, _ = timed(lambda: traj_euler_device(plan, z0, v0, B, dt, n_steps, 0.015, stream=stream,
                                                    schedule="grid"))
    t_adj, o = timed(lambda: trajectory_vjp(prob, theta, z0, v0, mol.mass, dt, n_steps - 1, ones, None, 0.015,
                                            None, 500, stream, plan=plan))
    t_tan, s = timed(lambda: trajectory_tangent(prob, theta, z0, v0, mol.mass, dt, n_steps - 1, 0.015, None,
                                                stream, plan=plan))
    t_tan_grid, _ = timed(lambda: trajectory_tangent(prob, theta, z0, v0, mol.mass, dt, n_steps - 1, 0.015, None,
                                                     stream, schedule="grid", plan=plan))
    # the generic kernels (tables read from device memory) on the same
    # workload, for comparison: plans built with specialisation off
    from molmot_amd.hip import plan as planmod
    saved = dict(planmod.SPECIALIZE_MIN_WORK)
    try:
        planmod.SPECIALIZE_MIN_WORK.update({k: float("inf") for k in saved})
        gplan = prob.plan(theta)
        t_fwd_gen, _ = timed(lambda: traj_euler_device(gplan, z0, v0, B, dt, n_steps, 0.015, stream=stream))
        t_tan_gen, _ = timed(lambda: trajectory_tangent(prob, theta, z0, v0, mol.mass, dt, n_steps - 1, 0.015, None,
                                                        stream, plan=gplan))
        t_adj_gen, _ = timed(lambda: trajectory_vjp(prob, theta, z0, v0, mol.mass, dt, n_steps - 1, ones, None,
                                                    0.015, None, 500, stream, plan=gplan))
    finally:
        planmod.SPECIALIZE_MIN_WORK.update(saved)
    g = o["g_theta"].sum(0).cpu().numpy()
    gt, gz0, gv0 = vjp_from_sens(s["sens"], ones, None)
    gt = gt.sum(0).cpu().numpy()
    rel = float(np.max(np.abs(gt - g)) / np.max(np.abs(g)))
    rel0 = float(max((gz0 - o["g_z0"]).abs().max() / o["g_z0"].abs().max(),
                     (gv0 - o["g_v0"]).abs().max() / o["g_v0"].abs().max()))
    return {"workload": f"{n} x {n_steps - 1} steps, no kicks; gradient of sum(z_final) w.r.t. "
                        "[det0..3, split, s0, B'] and (z0, v0)",
            "forward_s": t_fwd, "unit": "particle-steps/s",
            "timing": f"median of {repeats} runs, host wall time of the call; the forward and the gradients run on "
                      "one prebuilt plan (the plan argument of trajectory_tangent / trajectory_vjp)",
            "schedule": "auto (work queue at this size; the forward and the forward sensitivities alike)",
            "one_thread_per_particle": {"forward_s": t_fwd_grid, "forward_sensitivities_s": t_tan_grid,
                                        "gradient_cost_vs_forward": t_tan_grid / t_fwd_grid},
            "checkpointed_adjoint": {"seconds": t_adj, "gradient_cost_vs_forward": t_adj / t_fwd,
                                     "value": n * (n_steps - 1) / t_adj, "checkpoint_every": 500,
                                     "workspace_GB": 8 * (n * (2 * ((n_steps - 2) // 500 + 2) + 9 * 500 + 2)) / 1e9},
            "forward_sensitivities": {"seconds": t_tan, "gradient_cost_vs_forward": t_tan / t_fwd,
                                      "value": n * (n_steps - 1) / t_tan, "workspace_GB": 0.0,
                                      "max_rel_diff_vs_adjoint": {"theta": rel, "z0_v0": rel0}},
            "generic_kernels": {"forward_s": t_fwd_gen, "forward_sensitivities_s": t_tan_gen,
                                "checkpointed_adjoint_s": t_adj_gen,
                                "gradient_cost_vs_forward": t_tan_gen / t_fwd_gen},
            "kernels": "plan-specialised forward, forward sensitivities and checkpointed adjoint (hipRTC)",
            "gradient_cost_vs_forward": min(t_adj, t_tan) / t_fwd,
            "dsumz_dtheta": [float(x) for x in g]}


def size_scan(mol, plan, B, dt, ks, stream, n_steps=6000):
    """The ensemble at other sizes (6,000 steps each), with the schedule auto
    picks and with one thread per particle beside it.  One thread per
    particle and 64 per wavefront leave 100,000 particles as 1,563
    wavefronts on 1,024 SIMDs (half the SIMDs carry two, the rest one), and
    more than three per SIMD do not fit at once; the work queue evens that
    out.  Multiples of 65,536 particles up to 196,608 load every SIMD evenly,
    and there auto keeps one thread per particle."""
    import torch
    from molmot_amd import configs
    from molmot_amd.hip import traj_euler_device, traj_euler_schedule
    out = {"n_steps": n_steps, "unit": "particle-steps/s", "points": []}
    for n in (65536, 100000, 131072, 262144, 400000):
        z0_h, v0_h = configs.ensemble_initial_conditions(mol, n)
        z0, v0 = torch.from_numpy(z0_h).to("cuda"), torch.from_numpy(v0_h).to("cuda")
        pt = {"particles": n, "wavefronts_per_simd": n / 64 / 1024,
              "schedule": traj_euler_schedule(plan, n, n_steps, kick_scale=ks)}
        for sch in ("auto", "grid"):
            traj_euler_device(plan, z0, v0, B, dt, 300, 0.015, kick_scale=ks, seed=1, stream=stream, schedule=sch)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            traj_euler_device(plan, z0, v0, B, dt, n_steps, 0.015, kick_scale=ks, seed=2, stream=stream,
                              schedule=sch)
            b.record(stream)
            torch.cuda.synchronize()
            ms = a.elapsed_time(b)
            if sch == "auto":
                pt.update({"ms": ms, "value": n * n_steps / (ms / 1e3)})
            else:
                pt["one_thread_per_particle"] = {"ms": ms, "value": n * n_steps / (ms / 1e3)}
        out["points"].append(pt)
    return out


def traj3d_bench(mol, plan, B, dt, n_steps, stream, plan_generic=None):
    """The other trajectory models on 100,000 particles: Monte-Carlo
    temperature statistics of the 1D ensemble (all 30,000 steps), and the 3D
    ensembles (r0 ~ N(0, 1 mm) per axis, v0 5 mK per axis plus the
    [-15, 15] m/s sweep along z):
    * per-axis model (simulate_trajectories_3d: three 1D rate-equation
      forces per step at z_eff = (-x/2, -y/2, z), Euler), all 30,000 steps;
    * full Zeeman-eigenbasis force (MOTSimulator3D: the 12 x 12 complex
      ground Zeeman eigenbasis at every point, six beams x four frequencies
      of 5 mm radius, B' = 21.4 G/cm): points/s on 262,144 points and the
      leapfrog of the same ensemble over all 30,000 steps (a 200-step sample
      first; the full run when that sample extrapolates to under a minute),
      in the spin-rotation basis (one lane per particle), and beside it the
      per-point Jacobi diagonalisation (16 lanes per point) on a 200-step
      sample."""
    import torch
    from molmot_amd import configs
    from molmot_amd.hip import _lib
    from molmot_amd.obe.beam_pairs import make_beam_pairs_dc4
    from molmot_amd.mot.simulator_3d import make_6beam_config, Force3DPlan, force3d_points_device, leapfrog3d_device
    n = 100_000
    r0h, v0h = configs.ensemble_initial_conditions_3d(mol, n)
    r0, v0 = torch.from_numpy(r0h).to("cuda"), torch.from_numpy(v0h).to("cuda")
    rf, vf = torch.empty_like(r0), torch.empty_like(v0)
    na = torch.empty(n, dtype=torch.int32, device="cuda")

    ws3_bytes = int(_lib.lib().mm_traj_axis3d_workspace_bytes(n, n_steps))
    ws3 = torch.empty(ws3_bytes, dtype=torch.uint8, device="cuda")

    def axis(ns, schedule=0, p=plan):
        rc = _lib.lib().mm_traj_axis3d_ws(p.handle, n, _lib.ptr(r0), _lib.ptr(v0), float(B), float(mol.mass),
                                          float(dt), int(ns), 0.015, None, None, _lib.ptr(rf), _lib.ptr(vf),
                                          _lib.ptr(na), schedule, _lib.ptr(ws3), ws3_bytes,
                                          _lib.stream_handle(stream))
        _lib.check(rc, "mm_traj_axis3d_ws")

    def timed(fn):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        fn()
        b.record(stream)
        b.synchronize()
        return a.elapsed_time(b)

    # the headline's plan runs the plan-specialised kernels; the generic ones
    # of the same tables are timed beside them
    axis(10)
    ms_axis = timed(lambda: axis(n_steps))          # auto: the work queue at this size
    fin_axis = (rf.clone(), vf.clone(), na.clone())
    axis(10, 1)
    ms_axis_grid = timed(lambda: axis(n_steps, 1))  # one lane quad per particle
    axis_generic = None
    if plan_generic is not None:
        axis(10, 0, plan_generic)
        ms_axis_gen = timed(lambda: axis(n_steps, 0, plan_generic))
        axis_generic = {"value": n * n_steps / (ms_axis_gen / 1e3), "ms": ms_axis_gen,
                        "speedup_of_specialised": ms_axis_gen / ms_axis,
                        "bit_identical": all(bool(torch.equal(x, y)) for x, y in zip(fin_axis, (rf, vf, na)))}
    beams3 = make_6beam_config(mol, make_beam_pairs_dc4(mol, -1.0, 0.5, 2.0), 5e-3)
    plan3 = Force3DPlan(mol, beams3)                      # spin-rotation basis (SrOH tables)
    plan3j = Force3DPlan(mol, beams3, basis="jacobi")
    rng = np.random.default_rng(7)
    npts = 262_144
    rp = torch.from_numpy(rng.uniform(-5e-3, 5e-3, (npts, 3))).to("cuda")
    vp = torch.from_numpy(rng.uniform(-5.0, 5.0, (npts, 3))).to("cuda")
    z3 = {}
    for tag, pl in (("spin", plan3), ("jacobi", plan3j)):
        force3d_points_device(pl, rp[:1024], vp[:1024], B, stream=stream)
        z3[tag] = {"points_ms": timed(lambda: force3d_points_device(pl, rp, vp, B, stream=stream))}
        leapfrog3d_device(pl, r0[:256], v0[:256], B, dt, 2, save_full=False, stream=stream)
        z3[tag]["sample_ms"] = timed(lambda: leapfrog3d_device(pl, r0, v0, B, dt, 200, save_full=False,
                                                               stream=stream))
    full = None
    if z3["spin"]["sample_ms"] * n_steps / 200 < 60e3:
        out3 = {}

        def run_full():
            out3.update(leapfrog3d_device(plan3, r0, v0, B, dt, n_steps, save_full=False, stream=stream))

        full = timed(run_full)
        n_alive = int((out3["n_actual"] == n_steps).sum().item())
        n_eval = int(out3["n_actual"].to(torch.int64).sum().item())
    plan3.close()
    plan3j.close()
    lf = ({"workload": f"{n} particles x {n_steps} leapfrog steps (the full run)", "steps": n_steps,
           "value": n * n_steps / (full / 1e3), "unit": "particle-steps/s", "ms": full,
           "particles_inside_15mm_at_end": n_alive,
           "evaluated_particle_steps": n_eval,
           "note": "value counts every particle's 30,000 steps; escaped particles stop (evaluated_particle_steps). "
                   "The run goes in 256-step segments: after each, the particles still inside 15 mm are compacted "
                   "and continue with 1, 4 or 16 lanes each (mm_force3d_leapfrog_ws); one launch over all "
                   "particles takes 6.1 s, held up by wavefronts with one or two live lanes"}
          if full is not None else
          {"workload": f"{n} particles x 200 leapfrog steps (sample of the {n_steps}-step run)", "steps": 200,
           "value": n * 200 / (z3["spin"]["sample_ms"] / 1e3), "unit": "particle-steps/s",
           "ms": z3["spin"]["sample_ms"]})
    # ensemble moments at 50 sample steps (mm_traj_stats_1d: the device
    # sigma / temperature / density vs time), recoil kicks on
    from molmot_amd.hip import recoil_kick_scale
    zsh, vsh = configs.ensemble_initial_conditions(mol, n)
    zs_, vs_ = torch.from_numpy(zsh).to("cuda"), torch.from_numpy(vsh).to("cuda")
    bins = 50
    samp = torch.from_numpy(np.linspace(0, n_steps - 1, bins).astype(np.int32)).to("cuda")
    st_out = torch.empty((bins, 5), dtype=torch.float64, device="cuda")
    st_wb = int(_lib.lib().mm_traj_stats_workspace_bytes(n, bins))
    st_ws = torch.empty(max(st_wb, 8), dtype=torch.uint8, device="cuda")
    ks = recoil_kick_scale(mol)

    def stats(ns, p=plan):
        rc = _lib.lib().mm_traj_stats_1d(p.handle, n, _lib.ptr(zs_), _lib.ptr(vs_), float(B), float(mol.mass),
                                         float(dt), int(ns), 0.015, 0.015, _lib.ptr(samp), bins, _lib.ptr(st_out),
                                         ks, 3, 0, 0, _lib.ptr(st_ws), st_wb, _lib.stream_handle(stream))
        _lib.check(rc, "mm_traj_stats_1d")

    stats(n_steps)
    ms_stats = timed(lambda: stats(n_steps))
    st_spec = st_out.clone()
    stats_generic = None
    if plan_generic is not None:
        stats(n_steps, plan_generic)
        ms_stats_gen = timed(lambda: stats(n_steps, plan_generic))
        stats_generic = {"value": n * n_steps / (ms_stats_gen / 1e3), "ms": ms_stats_gen,
                         "speedup_of_specialised": ms_stats_gen / ms_stats,
                         "bit_identical": bool(torch.equal(st_out, st_spec))}
    ensemble_stats = {"workload": f"{n} particles x {n_steps} Euler steps with recoil kicks, (count, mean, M2) of "
                                  f"z and v over the molecules inside 15 mm at {bins} sample steps, one launch",
                      "value": n * n_steps / (ms_stats / 1e3), "unit": "particle-steps/s", "ms": ms_stats,
                      "kernel": "mms_traj_stats_k1 (plan-specialised)" if plan.specialized else "k_traj_stats",
                      "generic_kernel": stats_generic}
    # Monte-Carlo temperature statistics (k_traj_vsq): the 1D ensemble, v^2
    # summed over the second half of the run
    z1h, v1h = configs.ensemble_initial_conditions(mol, n)
    z1, v1 = torch.from_numpy(z1h).to("cuda"), torch.from_numpy(v1h).to("cuda")
    vsq = torch.empty(n, dtype=torch.float64, device="cuda")
    cnt = torch.empty(n, dtype=torch.int64, device="cuda")

    ws_bytes = int(_lib.lib().mm_traj_vsq_workspace_bytes(n, n_steps))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device="cuda")

    def mct(ns, schedule, p=plan):
        rc = _lib.lib().mm_traj_vsq_1d_ws(p.handle, n, _lib.ptr(z1), _lib.ptr(v1), float(B), float(mol.mass),
                                          float(dt), int(ns), int(ns) // 2, 0.015, _lib.ptr(vsq), _lib.ptr(cnt),
                                          schedule, _lib.ptr(ws), ws_bytes, _lib.stream_handle(stream))
        _lib.check(rc, "mm_traj_vsq_1d_ws")

    # the headline's plan runs the plan-specialised kernels; the generic
    # kernels (tables in memory) of the same tables are timed beside them
    mct(10, 0)
    ms_mct = timed(lambda: mct(n_steps, 0))   # auto: the work queue at this size
    sums = (vsq.clone(), cnt.clone())
    mct(10, 1)
    ms_mct_grid = timed(lambda: mct(n_steps, 1))
    generic = None
    if plan_generic is not None:
        mct(10, 0, plan_generic)
        ms_gen = timed(lambda: mct(n_steps, 0, plan_generic))
        generic = {"value": n * n_steps / (ms_gen / 1e3), "ms": ms_gen, "speedup_of_specialised": ms_gen / ms_mct,
                   "bit_identical": bool(torch.equal(vsq, sums[0]) and torch.equal(cnt, sums[1]))}
    return {"mc_temperature": {"workload": f"{n} particles x {n_steps} Euler steps, <v^2> over the second half "
                                           "(the 1D ensemble)",
                               "value": n * n_steps / (ms_mct / 1e3), "unit": "particle-steps/s", "ms": ms_mct,
                               "schedule": "auto (work queue at this size)",
                               "kernel": "mms_vsq_queue (plan-specialised)" if plan.specialized else "k_traj_vsq_queue",
                               "one_thread_per_particle": {"value": n * n_steps / (ms_mct_grid / 1e3),
                                                           "ms": ms_mct_grid},
                               "generic_kernel": generic},
            "ensemble_statistics": ensemble_stats,
            "per_axis": {"workload": f"{n} particles x {n_steps} Euler steps, 3 x 1D rate-equation force per step",
                         "value": n * n_steps / (ms_axis / 1e3), "unit": "particle-steps/s", "ms": ms_axis,
                         "schedule": "auto (work queue at this size: 6,250 wavefronts of 16 particles)",
                         "kernel": ("mms_axis3d_queue (plan-specialised)" if plan.specialized
                                    else "k_traj_axis3d_queue"),
                         "one_quad_per_particle": {"value": n * n_steps / (ms_axis_grid / 1e3), "ms": ms_axis_grid},
                         "generic_kernel": axis_generic},
            "zeeman_eigenbasis": {
                "basis": "spin rotation of the Zz eigenbasis (closed form, one lane per particle); the SrOH "
                         "tables' ground Zeeman blocks are spin-1/2 operators to 2e-12",
                "points": {"workload": f"{npts} (r, v) points, |r_i| < 5 mm, |v_i| < 5 m/s",
                           "value": npts / (z3["spin"]["points_ms"] / 1e3), "unit": "force-pts/s",
                           "ms": z3["spin"]["points_ms"]},
                "leapfrog": lf,
                "jacobi_basis": {
                    "points": {"value": npts / (z3["jacobi"]["points_ms"] / 1e3), "unit": "force-pts/s",
                               "ms": z3["jacobi"]["points_ms"]},
                    "leapfrog_200_steps": {"value": n * 200 / (z3["jacobi"]["sample_ms"] / 1e3),
                                           "unit": "particle-steps/s", "ms": z3["jacobi"]["sample_ms"]}}}}


def lindblad_flop_per_step(n):
    """RK4 step of one N-level density matrix: 4 right-hand sides, each one
    complex N x N x N product H rho (8 N^3) plus ~10 N^2 for the
    commutator's Hermitian part, dissipator and stage update."""
    return 4 * (8 * n ** 3 + 10 * n ** 2)


def lindblad_bench(mol, dev, stream, n_steps):
    """Standing-wave Lindblad force: 4096 velocities in [-2, 2] m/s x 64
    phases of the SrOH 12+4 class (four beam pairs, Omega = 1 Gamma, seeded
    random polarisation vectors, z = 1 mm in 0.2 T/m) and 4096 instances of a
    seeded 36+4 class, n_steps RK4 steps of 0.05/Gamma each (the full
    200-optical-period run is 25,133 steps; throughput is per
    instance-step).  FP64 VALU, FP32 VALU and FP32 MFMA paths; the error of
    each FP32 path is stated per force point against the FP64 path."""
    import torch
    from molmot_amd.molecules import synthetic_molecule
    from molmot_amd.obe.lindblad import LindbladClass, lindblad_rk4_device, steady_state_device
    G = mol.Gamma
    pairs = [{"freq": mol.omega_J32 + (-1.0 + 0.7) * G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0},
             {"freq": mol.omega_J32 + (-1.0 - 0.7) * G / (2 * np.pi), "q_fwd": 0, "q_bwd": 2, "s0": 1.0},
             {"freq": mol.omega_J12 + (-1.0 + 0.7) * G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0},
             {"freq": mol.omega_J12 + (-1.0 - 0.7) * G / (2 * np.pi), "q_fwd": 0, "q_bwd": 2, "s0": 1.0}]
    rng = np.random.default_rng(7)
    pol = rng.normal(size=(4, 2, 3)) + 1j * rng.normal(size=(4, 2, 3))
    pol /= np.linalg.norm(pol, axis=-1, keepdims=True)
    m40 = synthetic_molecule(36, 4, seed=11)
    Ee = np.mean(m40.energies[36:])
    pairs40 = [{"freq": Ee - np.mean(m40.energies[:12]) - G / (2 * np.pi), "q_fwd": 2, "q_bwd": 0, "s0": 1.0},
               {"freq": Ee - np.mean(m40.energies[12:36]) - G / (2 * np.pi), "q_fwd": 0, "q_bwd": 2, "s0": 1.0}]
    pol40 = rng.normal(size=(2, 2, 3)) + 1j * rng.normal(size=(2, 2, 3))
    pol40 /= np.linalg.norm(pol40, axis=-1, keepdims=True)
    out = {"workload": f"standing-wave Lindblad RK4, {n_steps} steps of dt=0.05/Gamma, Omega=1 Gamma per beam, "
                       "seeded random polarisations; 16-level: 4096 v in [-2,2] m/s x 64 phases; "
                       "40-level: 4096 (v, phase) instances",
           "unit": "instance-steps/s"}
    cases = [("sroh_12+4", mol, pairs, pol, 4096, 64, ("fp64", "fp32", "fp32_mfma")),
             ("synthetic_36+4", m40, pairs40, pol40, 64, 64, ("fp64", "fp32"))]
    from molmot_amd.autotune import Tuner, tune_lindblad_lanes
    tuner = Tuner(cache_path=None)   # time the candidates in this run
    for name, m, pr, pl, nv, nphi, precs in cases:
        cls = LindbladClass(m, pr, 1e-3, 0.2, 1.0, 1.0, pl)
        v = np.linspace(-2.0, 2.0, nv)
        V, P = np.meshgrid(v, np.linspace(0, 2 * np.pi, nphi, endpoint=False), indexing="ij")
        kv = torch.from_numpy((V * cls.kfac).ravel()).to(dev)
        ph = torch.from_numpy(P.ravel()).to(dev)
        n = kv.numel()
        res = {"instances": n, "n_levels": cls.N, "lanes_per_instance": {}}
        # lanes per instance of the VALU paths, picked by the autotuner on the
        # whole batch (10 steps per candidate)
        for prec in precs:
            if prec in ("fp64", "fp32"):
                best, ms = tune_lindblad_lanes(cls, prec, kv, ph, 0.05, 10, tuner=tuner)
                res["lanes_per_instance"][prec] = {"chosen": best, "ms_by_lanes": {str(k): v for k, v in ms.items()}}
        F64 = None
        for prec in precs:
            lindblad_rk4_device(cls, kv[:256], ph[:256], 0.05, 4, 2, prec, stream=stream)
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            F, R, _ = lindblad_rk4_device(cls, kv, ph, 0.05, n_steps, n_steps // 2, prec, stream=stream)
            b.record(stream)
            torch.cuda.synchronize()
            ms = a.elapsed_time(b)
            Fv = F.cpu().numpy().reshape(nv, nphi).mean(1)
            entry = {"avg_launch_ms": ms, "value": n * n_steps / (ms / 1e3),
                     "achieved_tflops": lindblad_flop_per_step(cls.N) * n * n_steps / (ms / 1e3) / 1e12}
            if prec == "fp64":
                F64 = Fv
            else:
                scale = float(np.max(np.abs(F64)))
                err = np.abs(Fv - F64) / scale
                entry["force_err_vs_fp64"] = {"max_rel": float(err.max()), "median_rel": float(np.median(err)),
                                              "normalised_by": "max |F_fp64(v)| over the velocity grid"}
                entry["speedup_vs_fp64"] = res["fp64"]["avg_launch_ms"] / ms
            res[prec] = entry
        out[name] = res
    # uneven batch in one launch at the configuration's size: 4096 v x 64 phases of each class
    from molmot_amd.obe.lindblad import lindblad_rk4_uneven_device
    c16 = LindbladClass(mol, pairs, 1e-3, 0.2, 1.0, 1.0, pol)
    c40 = LindbladClass(m40, pairs40, 1e-3, 0.2, 1.0, 1.0, pol40)
    nvu, npu = 4096, 64
    V, P = np.meshgrid(np.linspace(-2.0, 2.0, nvu), np.linspace(0, 2 * np.pi, npu, endpoint=False), indexing="ij")
    kv16 = torch.from_numpy((V * c16.kfac).ravel()).to(dev)
    kv40 = torch.from_numpy((V * c40.kfac).ravel()).to(dev)
    ph = torch.from_numpy(P.ravel()).to(dev)
    n1 = kv16.numel()
    flop = (lindblad_flop_per_step(16) + lindblad_flop_per_step(40)) * n1 * n_steps
    un = {"workload": f"{n1} 12+4 (SrOH) + {n1} 36+4 (synthetic) instances ({nvu} v x {npu} phases per class) in "
                      f"one launch, {n_steps} RK4 steps (sample of the 25,133-step run); 16-level: one wavefront "
                      "each (MFMA in FP32), 40-level: four wavefronts each",
          "unit": "instance-steps/s"}
    ref = None
    for prec in ("fp64", "fp32_mfma"):
        lindblad_rk4_uneven_device(c16, kv16[:64], ph[:64], c40, kv40[:64], ph[:64], 0.05, 4, 2, prec)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        (F16, _, _), (F40, _, _) = lindblad_rk4_uneven_device(c16, kv16, ph, c40, kv40, ph, 0.05, n_steps,
                                                              n_steps // 2, prec, stream=stream)
        b.record(stream)
        torch.cuda.synchronize()
        ms = a.elapsed_time(b)
        e = {"avg_launch_ms": ms, "value": 2 * n1 * n_steps / (ms / 1e3),
             "achieved_tflops": flop / (ms / 1e3) / 1e12}
        Fs = (F16.cpu().numpy().reshape(nvu, npu).mean(1), F40.cpu().numpy().reshape(nvu, npu).mean(1))
        if prec == "fp64":
            ref = Fs
        else:
            e["force_err_vs_fp64"] = {
                k: float(np.max(np.abs(f - r)) / np.max(np.abs(r))) for k, f, r in zip(("16", "40"), Fs, ref)}
            e["speedup_vs_fp64"] = un["fp64"]["avg_launch_ms"] / ms
        un[prec] = e
    out["uneven_16_40"] = un
    # steady-state drop-in (reference call): 16 phases of the 12+4 class
    cls = LindbladClass(mol, pairs, 1e-3, 0.2, 1.0, 1.0, pol)
    ph = torch.from_numpy(np.linspace(0, 2 * np.pi, 16, endpoint=False)).to(dev)
    steady_state_device(cls, ph, clip_negative=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        steady_state_device(cls, ph, clip_negative=False)
    torch.cuda.synchronize()
    out["steady_state_12+4_16_phases_ms"] = (time.perf_counter() - t0) / 5 * 1e3
    return out


def floquet_bench(mol, dev, stream, n_v=512, n_max=3):
    """Floquet sub-Doppler force (molmot/obe/floquet.py:61-403) for SrOH
    12+4 with the 4-frequency DC beam pairs (delta -1 Gamma, split 0.5
    Gamma, s0 2, Omega 1 Gamma): n_v velocities in [-2, 2] m/s solved as one
    batch, n_max = 3 (7 Fourier blocks of 256 x 256); the CPU leg checks two
    of the forces against the FP64 oracle's dense solve."""
    import torch
    from molmot_amd.obe.beam_pairs import make_beam_pairs_dc4
    from molmot_amd.obe.floquet import FloquetOBE
    pairs = make_beam_pairs_dc4(mol, -1.0, 0.5, 2.0)
    fo = FloquetOBE(mol, pairs, 0.2, 1.0, n_max, 1.0)
    v = np.linspace(-2.0, 2.0, n_v)
    fo.force_vs_velocity(v[:8])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    F = fo.force_vs_velocity(v)
    el = time.perf_counter() - t0
    res = {"workload": f"SrOH 12+4, DC 4-frequency pairs, {n_v} velocities in [-2, 2] m/s, n_max = {n_max} "
                       f"({2 * n_max + 1} blocks of 256 x 256 complex)",
           "value": n_v / el, "unit": "velocities/s", "seconds": el}
    return res, (pairs, v, F, n_max)


def floquet_parity(mol, check):
    """The Floquet forces' error on two velocities against the FP64
    oracle's dense solve of the full system (CPU leg)."""
    import oracle.floquet as ofq
    pairs, v, F, n_max = check
    idx = [v.size // 4, (3 * v.size) // 4]
    Fo = ofq.force(mol, ofq.blocks(mol, pairs, 1.0, 1.0), v[idx], n_max)
    return {"max_abs_over_max_F": float(np.max(np.abs(F[idx] - Fo)) / np.max(np.abs(F))),
            "sample": "2 velocities, dense LAPACK solve of the full system"}


# SrOH MCWF step, the reference algorithm (sse_trajectory_kernel,
# molmot_cuda/stochastic_cuda.py:404-756): 4 x 7.3 kflop + 0.3 kflop
MCWF_FLOP_PER_STEP_REF = 29500


def mcwf_flop_per_step(ng, ne):
    """Algorithmic flops of one RK4 step of one trajectory in the MCWF kernel
    (FMA = 2): per right-hand side, interaction phases (N complex products),
    six-beam fields (18 complex products + sums), the one-pass dipole
    contraction (per ground state 3 n_e (2+2) FMA + d_exp + coupling +
    back-transform), the excited rows, the Hermitian Zeeman block (66 upper
    entries x (6 FMA to build H + 8 FMA to apply it both ways)) and the force;
    plus the RK4 combination."""
    n = ng + ne
    rhs = 6 * ne + 18 * 6 + 9 * 8 + ng * (6 + 3 * ne * 8 + 3 * 8 + 3 * 8 + 8) + ne * (3 * 8 + 8 + 6) \
        + (ng * (ng - 1) // 2) * 28 + 36
    return 4 * rhs + 3 * 8 * n + 6 * n


def mcwf_bench(mol, dev, stream, n_steps):
    """MCWF ensemble: 65,536 SrOH wavefunctions (three groups starting at
    v_z = 0, 0.5, 1 m/s at the trap centre), four-frequency DC beams
    (delta = -1 Gamma, split 0.5 Gamma, s0 = 2 per beam, 10 mm radius),
    B' = 0.16 T/m, dt = 0.01/Gamma, recoil kicks on, FP32 amplitudes with FP64
    norm, jump threshold, r and v, exact waiting-time jumps.  Reports trajectory-steps/s, the
    ensemble force, momentum diffusion and photon scattering rate per group,
    and an FP64-amplitude throughput on a 2,000-step sample.  Flops per step:
    the reference algorithm's count (4 RHS x 7.3 kflop + 0.3 kflop, with ~92
    sincos not counted) as the algorithmic figure, this kernel's own
    structured count beside it.

    For the comparison with the FP64 Lindblad force it also runs, at each of
    v = 0, 0.5, 1 m/s, 16,384 trajectories of the same problem at fixed
    velocity (mass x 1e12, no recoil) for 4,000 steps, averaging each
    trajectory's normalised force over steps 1,000-3,999; the returned
    check data lets the CPU leg integrate the master equation this MCWF
    unravels over the same steps."""
    import dataclasses
    import torch
    from molmot_amd.obe.beam_pairs import make_beam_pairs_dc4
    from molmot_amd.obe.stochastic import SSEProblem, MCWFPlan, mcwf_ensemble_device
    from molmot_amd.obe.diffusion import compute_diffusion_from_ensemble_device
    bp = make_beam_pairs_dc4(mol, -1.0, 0.5, 2.0)
    B = 0.16
    prob = SSEProblem(mol, bp, B)
    plan = MCWFPlan(prob)
    n = 65536
    groups = (0.0, 0.5, 1.0)
    v0 = np.zeros((n, 3))
    gid = np.arange(n) % 3
    v0[:, 2] = np.array(groups)[gid] / prob.v_nat
    r0 = torch.zeros((n, 3), dtype=torch.float64, device=dev)
    v0t = torch.from_numpy(v0).to(dev)
    save_every = max(1, n_steps // 50)
    mcwf_ensemble_device(plan, r0, v0t, 0.01, 10, 10, 1, precision="fp32", want=(), stream=stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = mcwf_ensemble_device(plan, r0, v0t, 0.01, n_steps, save_every, 7, precision="fp32",
                               avg_start=n_steps // 5, jump_rule="exact", want=("v", "force_sum"), stream=stream)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    fl = MCWF_FLOP_PER_STEP_REF
    res = {"workload": f"{n} SrOH 12+4 wavefunctions x {n_steps} RK4 steps, dt=0.01/Gamma, 6 beams x 4 freq, "
                       "recoil on, FP32 amplitudes / FP64 norm, jumps, r, v; exact waiting-time jump rule "
                       "(the reference rule 1-|psi|^2 > -ln U strands trajectories whose threshold exceeds 1)",
           "value": n * n_steps / el, "unit": "trajectory-steps/s", "seconds": el,
           "flop_per_unit": fl, "achieved_tflops": fl * n * n_steps / el / 1e12,
           "flop_per_unit_note": "the reference algorithm's count per trajectory-step (4 RHS x 7.3 kflop "
                                 "+ 0.3 kflop; ~92 sincos not counted)",
           "executed_flop_per_unit_est": mcwf_flop_per_step(mol.n_ground, mol.n_excited)}
    g = torch.from_numpy(gid).to(dev)
    t_s = out["times"] * prob.t_nat
    fs = out["force_sum"] / out["n_avg"].clamp(min=1).to(torch.float64)[:, None]
    per = {}
    for i, vg in enumerate(groups):
        m = g == i
        F = fs[m, 2] * prob.F_nat
        D, D_err, _ = compute_diffusion_from_ensemble_device(out["v"][m] * prob.v_nat, t_s, prob.mass, axis=0)
        R = out["n_scatters"][m].double().mean().item() / (n_steps * 0.01 * prob.t_nat)
        per[f"v0={vg}"] = {"F_z_mean_N": F.mean().item(), "F_z_stderr_N": (F.std() / m.sum().sqrt()).item(),
                           "D_x_kg2m2s3": D, "D_x_err": D_err, "scatter_rate_per_s": R}
    res["estimators"] = per
    # fixed-velocity ensembles for the master-equation comparison
    heavy = dataclasses.replace(mol, mass=mol.mass * 1e12)
    bph = make_beam_pairs_dc4(heavy, -1.0, 0.5, 2.0)
    probh = SSEProblem(heavy, bph, B, add_spontaneous_kick=False)
    planh = MCWFPlan(probh)
    nm, sm, am = 16384, 4000, 1000
    fixed = []
    for i, vg in enumerate(groups):
        vv = torch.zeros((nm, 3), dtype=torch.float64, device=dev)
        vv[:, 2] = vg / probh.v_nat
        o = mcwf_ensemble_device(planh, torch.zeros((nm, 3), dtype=torch.float64, device=dev), vv, 0.01, sm, sm,
                                 31 + i, precision="fp32", avg_start=am, jump_rule="exact",
                                 want=("force_sum",), stream=stream)
        fz = (o["force_sum"][:, 2] / o["n_avg"].clamp(min=1).to(torch.float64)) * probh.F_nat
        fixed.append((vg, fz.mean().item(), (fz.std() / nm ** 0.5).item()))
    torch.cuda.synchronize()
    check = (heavy, bph, B, planh.shift, probh.v_nat, probh.F_nat, sm, am, fixed)
    ns64 = min(2000, n_steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mcwf_ensemble_device(plan, r0, v0t, 0.01, ns64, ns64, 7, precision="fp64", want=(), stream=stream)
    torch.cuda.synchronize()
    el64 = time.perf_counter() - t0
    res["fp64"] = {"value": n * ns64 / el64, "unit": "trajectory-steps/s", "steps": ns64}
    return res, check


def mcwf_master_parity(check):
    """The fixed-velocity ensemble forces against the FP64 master equation
    of the same problem (oracle.mcwf.lindblad_path: the Lindblad equation
    this MCWF unravels, from the same initial ground state), time-averaged
    over the same steps (CPU leg)."""
    import oracle.mcwf as omc
    heavy, bp, B, shift, v_nat, F_nat, n_steps, avg_start, fixed = check
    tab = omc.tables(heavy, bp, B, add_spontaneous_kick=False)
    rho0 = np.zeros((heavy.n_states, heavy.n_states), dtype=complex)
    rho0[0, 0] = 1.0
    rows = []
    for vg, Fm, se in fixed:
        _, F, _ = omc.lindblad_path(tab, np.zeros(3), np.array([0.0, 0.0, vg / v_nat]), rho0, 0.01, n_steps, shift)
        Fl = float(np.mean(F[avg_start:n_steps, 2])) * F_nat
        rows.append({"v_ms": vg, "F_mcwf_N": Fm, "F_mcwf_stderr_N": se, "F_lindblad_fp64_N": Fl,
                     "dev_over_stderr": abs(Fm - Fl) / se if se > 0 else None})
    return {"sample": f"16384 trajectories per velocity at fixed v (no recoil), {n_steps} steps of 0.01/Gamma, "
                      f"force averaged over steps {avg_start}-{n_steps - 1}; FP64 master equation of the same "
                      "six-beam problem over the same steps",
            "points": rows}


def rf_mot_bench(mol, z0, v0, dt, n_steps):
    """The RF MOT's 1D ensemble (molmot/mot/simulator.py:49-90: every step
    averages the phase-0 force at +B' and the phase-1 force at -B'; SrOH,
    detuning -1 Gamma, s0 = 1, B' = 0.16 T/m) on the headline's initial
    conditions, all steps, no kicks: the plan-specialised RF kernels
    (mms_rf_traj_k0, both phase plans as constants) against the generic RF
    kernel (k_traj_euler with RF, tables in memory); events on the launch
    stream, best of 3 after a warm-up."""
    import torch
    from molmot_amd.constants import hbar as HBAR
    from molmot_amd.hip.plan import HIPMolData
    from molmot_amd.hip.trajectories import traj_euler_rf_device
    from molmot_amd.obe.fields import make_rf_mot_beams
    p0, p1 = (HIPMolData(mol, make_rf_mot_beams(mol, delta_Gamma=-1.0, s0=1.0, phase=ph), hbar=HBAR,
                         specialize=False) for ph in (0, 1))
    n = z0.numel()
    out, res = {}, {}
    for mode in ("generic", "specialised"):
        p0.specialize_rf(p1 if mode == "specialised" else None)
        traj_euler_rf_device(p0, p1, z0, v0, 0.16, dt, n_steps)
        ts = []
        for _ in range(3):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            r = traj_euler_rf_device(p0, p1, z0, v0, 0.16, dt, n_steps)
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        res[mode] = r
        steps = float((r[2].to(torch.float64) - 1.0).sum().item())  # updates made: escaped particles stop
        out[mode] = {"ms": min(ts), "value": steps / (min(ts) / 1e3)}
    same = all(torch.equal(a, b) for a, b in zip(res["generic"][:3], res["specialised"][:3]))
    p0.specialize_rf(None)
    return {"workload": f"{n} particles x {n_steps} Euler steps, RF MOT (two switching phases per step), no kicks",
            "unit": "particle-steps/s (updates made, sum of n_actual - 1)",
            "kernel": "mms_rf_traj_k0 (plan-specialised RF pair)",
            "specialised": out["specialised"], "generic_kernel": out["generic"],
            "speedup_of_specialised": out["generic"]["ms"] / out["specialised"]["ms"],
            "escaped": int((res["specialised"][2] < n_steps).sum().item()), "bit_identical": bool(same)}


def spec_grid_timings():
    """HIP-event times of the plan-specialised kernels in their one-thread-
    per-particle forms, three launches each after a warm-up: the headline
    ensemble (100,000 x 30,000 steps, kicks) on mms_traj_grid_k1, the 1024^2
    force map on mms_rate_map and the ensemble statistics (50 samples,
    kicks) on mms_traj_stats_k1."""
    import torch
    from molmot_amd import configs
    from molmot_amd.constants import hbar
    from molmot_amd.hip import HIPMolData, _lib, traj_euler_device, recoil_kick_scale, rate_eq_map_device
    from molmot_amd.molecules import load_sroh_from_julia
    mol = load_sroh_from_julia()
    plan = HIPMolData(mol, configs.red_mot_beams(mol), 1.0, hbar=hbar, specialize=True)
    B, dt = configs.B_GRAD_SI, configs.ENSEMBLE_DT
    n, ns = configs.ENSEMBLE_N, configs.ENSEMBLE_STEPS
    zh, vh = configs.ensemble_initial_conditions(mol, n)
    z0, v0 = torch.from_numpy(zh).to("cuda"), torch.from_numpy(vh).to("cuda")
    ks = recoil_kick_scale(mol)
    stream = torch.cuda.current_stream()

    def times(fn, reps=3):
        fn()
        torch.cuda.synchronize()
        out = []
        for _ in range(reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            fn()
            b.record(stream)
            torch.cuda.synchronize()
            out.append(a.elapsed_time(b))
        return out

    res = {"ensemble_grid": {"kernel": "mms_traj_grid_k1", "workload": f"{n} x {ns} steps, kicks",
                             "ms": times(lambda: traj_euler_device(plan, z0, v0, B, dt, ns, 0.015, kick_scale=ks,
                                                                   seed=5, stream=stream, schedule="grid"))}}
    vg, zg = configs.force_map_grid()
    vt, zt = torch.from_numpy(vg).to("cuda"), torch.from_numpy(zg).to("cuda")
    bufs = rate_eq_map_device(plan, vt, zt, B)
    res["force_map"] = {"kernel": "mms_rate_map", "workload": "1024 x 1024 points, F + populations + R",
                        "ms": times(lambda: rate_eq_map_device(plan, vt, zt, B, out=bufs, stream=stream), 20)}
    bins = 50
    samp = torch.from_numpy(np.linspace(0, ns - 1, bins).astype(np.int32)).to("cuda")
    out = torch.empty((bins, 5), dtype=torch.float64, device="cuda")
    wb = int(_lib.lib().mm_traj_stats_workspace_bytes(n, bins))
    ws = torch.empty(max(wb, 8), dtype=torch.uint8, device="cuda")
    res["ensemble_statistics"] = {
        "kernel": "mms_traj_stats_k1", "workload": f"{n} x {ns} steps, kicks, {bins} samples",
        "ms": times(lambda: _lib.check(_lib.lib().mm_traj_stats_1d(
            plan.handle, n, _lib.ptr(z0), _lib.ptr(v0), float(B), float(mol.mass), float(dt), ns, 0.015, 0.015,
            _lib.ptr(samp), bins, _lib.ptr(out), ks, 3, 0, 0, _lib.ptr(ws), wb, _lib.stream_handle(stream))))}
    from molmot_amd.adjoint import ForceProblem, trajectory_tangent
    prob = ForceProblem(mol, configs.red_mot_beams(mol))
    theta = prob.theta0(B).numpy()
    gplan = prob.plan(theta)
    gplan.specialize(2)
    res["forward_sensitivities_grid"] = {
        "kernel": "mms_tangent_grid", "workload": f"{n} x {ns - 1} updates, no kicks",
        "ms": times(lambda: trajectory_tangent(prob, theta, z0, v0, mol.mass, dt, ns - 1, 0.015, None, stream,
                                               schedule="grid", check=False, plan=gplan))}
    res["forward_grid_no_kicks"] = {"kernel": "mms_traj_grid_k0", "workload": f"{n} x {ns} steps, no kicks",
                                    "ms": times(lambda: traj_euler_device(plan, z0, v0, B, dt, ns, 0.015,
                                                                          stream=stream, schedule="grid"))}
    for r in res.values():
        r["avg_ms"] = float(np.mean(r["ms"]))
    print(json.dumps({"spec_grid_event_times": res}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--particles", type=int, default=None)
    ap.add_argument("--n-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-map", action="store_true")
    ap.add_argument("--no-lindblad", action="store_true")
    ap.add_argument("--no-traj3d", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--lindblad-steps", type=int, default=200)
    ap.add_argument("--no-mcwf", action="store_true")
    ap.add_argument("--no-floquet", action="store_true")
    ap.add_argument("--mcwf-steps", type=int, default=50000)
    ap.add_argument("--spec-grid", action="store_true",
                    help="only time the plan-specialised kernels in their one-thread-per-particle forms "
                         "(headline ensemble, force map, ensemble statistics, forward sensitivities) with HIP "
                         "events, print one JSON "
                         "line and exit: the run a rocprofv3 kernel trace of those kernels is checked against")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy a 512 MiB device buffer three times (calibrates the FETCH_SIZE and "
                         "WRITE_SIZE counters against a known byte count) and exit")
    args = ap.parse_args()
    if args.spec_grid:
        spec_grid_timings()
        return
    if args.copy_probe:
        import torch
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        for _ in range(3):
            x.clone()
        torch.cuda.synchronize()
        print("copied 3 x 512 MiB", flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the ensemble shards by "
              "particle ranges across GPUs, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    import torch
    from molmot_amd import configs
    from molmot_amd.constants import hbar
    from molmot_amd.hip import HIPMolData, traj_euler_device, traj_euler_schedule, recoil_kick_scale, \
        rate_eq_map_device
    from molmot_amd.molecules import load_sroh_from_julia

    dev = torch.device("cuda", 0)
    mol = load_sroh_from_julia()
    beams = configs.red_mot_beams(mol)
    B = configs.B_GRAD_SI
    N = args.particles or configs.ENSEMBLE_N
    n_steps = args.n_steps or configs.ENSEMBLE_STEPS
    dt = configs.ENSEMBLE_DT
    z0_h, v0_h = configs.ensemble_initial_conditions(mol, N)
    # the plan-specialised kernels (tables compiled into the code, hipRTC);
    # the generic kernels of the same plan are timed beside them
    plan = HIPMolData(mol, beams, 1.0, hbar=hbar, specialize=False)
    t_c = time.perf_counter()
    plan.specialize()
    spec_compile_s = time.perf_counter() - t_c
    plan_generic = HIPMolData(mol, beams, 1.0, hbar=hbar, specialize=False)
    z0 = torch.from_numpy(z0_h).to(dev)
    v0 = torch.from_numpy(v0_h).to(dev)
    ks = recoil_kick_scale(mol)
    stream = torch.cuda.current_stream()

    def run(seed):
        return traj_euler_device(plan, z0, v0, B, dt, n_steps, 0.015, kick_scale=ks, seed=seed, stream=stream)

    schedule = traj_euler_schedule(plan, N, n_steps, kick_scale=ks)

    for w in range(args.warmup):
        run(1000 + w)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for s in range(args.steps):
        ev[s][0].record(stream)
        out = run(s)
        ev[s][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in ev]
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    ms_per_step = wall / args.steps * 1e3
    value = N * n_steps / (wall / args.steps)
    # the generic kernels (tables read from device memory) on the same workload
    gen_ms = []
    for s in range(args.steps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        gen_out = traj_euler_device(plan_generic, z0, v0, B, dt, n_steps, 0.015, kick_scale=ks, seed=s,
                                    stream=stream)
        b.record(stream)
        torch.cuda.synchronize()
        gen_ms.append(a.elapsed_time(b))
    gen_identical = all(torch.equal(x, y) for x, y in zip(out[:3], gen_out[:3]))
    na = out[2].cpu().numpy()
    # particle-steps evaluated in a launch: an escaped particle stops at its
    # escape step (n_actual), so sum(n_actual - 1) steps are integrated
    evaluated = float(np.sum(na.astype(np.int64) - 1))
    active = evaluated / float(N * (n_steps - 1))
    zf = out[0].cpu().numpy()
    captured = float(np.mean((na == n_steps) & (np.abs(zf) < 3e-3)))
    # executed work: the FP64 operations of one specialised step, counted in
    # the kernel's code object (2 x FMA + MUL + ADD; the step loop is
    # straight-line code), and the dense algorithm's count beside it
    kname = "mms_traj_queue_k1" if schedule == "queue" else "mms_traj_grid_k1"
    try:
        from molmot_amd.hip import spec as specmod
        counts = specmod.fp64_counts(specmod.compile_source(plan.spec_source()), kname)
    except Exception as exc:  # llvm-objdump absent: no executed count
        counts = None
        _progress(f"no FP64 instruction count: {exc}")
    exec_flop = counts["flop"] if counts else None
    achieved_tflops = exec_flop * evaluated / avg_kernel_s / 1e12 if exec_flop else None
    dense_tflops = FLOP_PER_PARTICLE_STEP * evaluated / avg_kernel_s / 1e12
    # the same initial conditions in the reference's DC MOT (delta = -0.20
    # Gamma, split 0.70 Gamma, s0 = 1, B' = 0.16 T/m), where they are trapped
    from molmot_amd.obe.fields import make_dc_mot_beams
    dc_plan = HIPMolData(mol, make_dc_mot_beams(mol, -0.2, 0.7, 1.0), 1.0, hbar=hbar)
    zdc, _, nadc, _, _ = traj_euler_device(dc_plan, z0, v0, 0.16, dt, n_steps, 0.015, kick_scale=ks, seed=7,
                                           stream=stream)
    nadc, zdc = nadc.cpu().numpy(), zdc.cpu().numpy()
    captured_dc = float(np.mean((nadc == n_steps) & (np.abs(zdc) < 3e-3)))
    _progress(f"ensemble: {value:.3e} particle-steps/s")

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"1D rate-equation trajectory ensemble: {N} SrOH (12+4) x {n_steps} "
                        f"symplectic-Euler steps, dt={dt:g} s, 4-freq red MOT (100 mW/6 beams, w=5 mm), "
                        "B'=21.4 G/cm, Philox recoil kicks, FP64",
            "particles": N, "n_steps": n_steps, "dt": dt,
            "parallelism": ("1 GPU, 1 thread per particle; work queue of (64-particle group, chunk of steps) "
                            "units on a persistent grid of 2 wavefronts per SIMD" if schedule == "queue"
                            else "1 GPU, 1 thread per particle"),
            "kernels": "plan-specialised (the plan's tables compiled into the code with hipRTC)",
        },
        "plan_specialised": {
            "compile_s": spec_compile_s,
            "generic_kernel": {"kernel": ("k_traj_euler_queue<12,4>" if schedule == "queue"
                                          else "k_traj_euler<12,4>"),
                               "avg_launch_ms": float(np.mean(gen_ms)),
                               "value": N * n_steps / (float(np.mean(gen_ms)) / 1e3),
                               # rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of the generic kernel at
                               # 100,000 x 30,000 (KiB per dispatch: 449,476.6 and 493,385.5),
                               # scaled by what the counters read for a 512 MiB device copy
                               # (x2.0 and x1.0): nearly all of it the queue's hand-over of each
                               # particle's state (z, v, n_actual: 20 B read and written per
                               # chunk, 235 chunks); one thread per particle: 1298.6 and
                               # 2734.4 KiB.  Other sizes: None (not measured)
                               "traffic": ((449476.6 * 2.0 + 493385.5 * 1.0) * 1024 if schedule == "queue"
                                           else (1298.6 * 2.0 + 2734.4 * 1.0) * 1024)
                               if (N, n_steps) == (100000, 30000) else None},
            "speedup_vs_generic": float(np.mean(gen_ms)) / (avg_kernel_s * 1e3),
            "bit_identical_to_generic": bool(gen_identical),
        },
        "roofline": {
            "bound": "fp64-valu",
            "pipe": "FP64 VALU (MI355X FP64 vector peak = matrix peak); no MFMA in this kernel",
            "kernel": kname,
            "achieved": achieved_tflops,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / PEAK_FP64_TFLOPS if achieved_tflops else None,
            # HBM traffic of the specialised kernel: not measured (null).  Its
            # queue hands each particle's state over between chunks as the
            # generic kernel's does; the generic kernel's counters are under
            # plan_specialised.generic_kernel.traffic
            "traffic": None,
            "traffic_algorithmic": 36 * N,
            "flop_per_unit": exec_flop,
            "flop_per_unit_source": ("FP64 VALU instructions of the kernel in its code object (hipRTC, disassembled): "
                                     "2 x FMA + MUL + ADD; the step loop is straight-line code"),
            "fp64_instructions_per_step": counts,
            "units_per_launch": evaluated,
            "units": "particle-steps evaluated in the launch, sum(n_actual - 1) (escaped particles stop)",
            "avg_launch_ms": avg_kernel_s * 1e3,
            "dense_equivalent": {"flop_per_unit": FLOP_PER_PARTICLE_STEP, "tflops": dense_tflops,
                                 "frac": dense_tflops / PEAK_FP64_TFLOPS,
                                 "note": "the reference algorithm's dense count per step (12 x 12 elimination); "
                                         "the reduced solve executes fewer operations"},
        },
        "ensemble": {"active_step_fraction": active, "capture_fraction_3mm": captured,
                     "capture_fraction_3mm_dc_mot": captured_dc,
                     "note": "the 4-frequency red MOT of this workload has a negative 1D spring constant at "
                             "B' = 21.4 G/cm, so its ensemble is not captured; capture_fraction_3mm_dc_mot is the "
                             "fraction of the same initial conditions (most start at |v| of several m/s) held "
                             "within 3 mm by the reference's DC MOT (delta -0.20 Gamma, split 0.70 Gamma, s0 1, "
                             "B' 0.16 T/m) at the end of the run"},
    }

    if not args.no_map:
        v_g, z_g = configs.force_map_grid()
        vt = torch.from_numpy(v_g).to(dev)
        zt = torch.from_numpy(z_g).to(dev)
        npts = v_g.size * z_g.size
        reps = 20
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

        def timed(launch):
            for _ in range(3):
                launch()
            torch.cuda.synchronize()
            a.record(stream)
            for _ in range(reps):
                launch()
            b.record(stream)
            torch.cuda.synchronize()
            return a.elapsed_time(b) / reps

        # the specialised plan runs the generated map kernel; the generic plan
        # the table-driven one, threads per block chosen live by the autotuner
        # (one thread per point; the block size sets the LDS slab per block)
        bufs = rate_eq_map_device(plan, vt, zt, B)
        torch.cuda.synchronize()
        map_out = (bufs[0].cpu().numpy().reshape(-1), bufs[1].cpu().numpy().reshape(npts, -1),
                   bufs[2].cpu().numpy().reshape(-1), v_g, z_g)
        map_ms = timed(lambda: rate_eq_map_device(plan, vt, zt, B, out=bufs, stream=stream))
        from molmot_amd.autotune import Tuner, tune_rate_block
        map_block, map_block_ms = tune_rate_block(plan_generic, vt, zt, B, tuner=Tuner(cache_path=None))
        gbufs = rate_eq_map_device(plan_generic, vt, zt, B)
        gen_map_ms = timed(lambda: rate_eq_map_device(plan_generic, vt, zt, B, out=gbufs, stream=stream))
        map_same = all(bool(torch.equal(x, y)) for x, y in zip(bufs, gbufs))
        del gbufs
        # adjoint: F and dF/d[4 detunings, split, s0, B', z, v] at every grid
        # point, specialised gradient kernel and generic one
        from molmot_amd.adjoint import ForceProblem
        from molmot_amd.hip.plan import GRADIENT
        from molmot_amd.hip import _lib
        prob = ForceProblem(mol, beams)
        theta = prob.theta0(B).numpy()
        gplan = prob.plan(theta)
        gplan_spec = prob.plan(theta)
        t_c = time.perf_counter()
        gplan_spec.specialize(GRADIENT)
        grad_compile_s = time.perf_counter() - t_c
        V, Z = torch.meshgrid(vt, zt, indexing="ij")
        Vf, Zf = V.contiguous().reshape(-1), Z.contiguous().reshape(-1)
        Fa = torch.empty(npts, dtype=torch.float64, device=dev)
        Ja = torch.empty((npts, 9), dtype=torch.float64, device=dev)
        Fg = torch.empty_like(Fa)
        Jg = torch.empty_like(Ja)

        def grad_points(p, F_, J_):
            _lib.check(_lib.lib().mm_rate_eq_grad_points(p.handle, npts, _lib.ptr(Vf), _lib.ptr(Zf), B,
                                                         _lib.ptr(F_), _lib.ptr(J_), _lib.stream_handle(stream)))

        adj_ms = timed(lambda: grad_points(gplan_spec, Fa, Ja))
        gen_adj_ms = timed(lambda: grad_points(gplan, Fg, Jg))
        ok = torch.isfinite(Ja).all(dim=1) & torch.isfinite(Jg).all(dim=1)
        jrel = float(((Ja - Jg).abs().amax(dim=1)[ok] / Jg.abs().amax(dim=1)[ok].clamp_min(1e-300)).max())
        f_same = bool(torch.equal(Fa, Fg))
        result["force_map"] = {"workload": "F(v,z) 1024x1024, FP64 forward, F + 12 populations + R",
                               "value": npts / (map_ms / 1e3), "unit": "force-pts/s",
                               "avg_launch_ms": map_ms, "kernel": "mms_rate_map (plan-specialised)",
                               "achieved_tflops": FLOP_PER_EVAL * npts / (map_ms / 1e3) / 1e12,
                               "generic_kernel": {"avg_launch_ms": gen_map_ms,
                                                  "value": npts / (gen_map_ms / 1e3),
                                                  "block": {"chosen": map_block, "ms_by_block": map_block_ms},
                                                  "speedup_of_specialised": gen_map_ms / map_ms,
                                                  "bit_identical": map_same},
                               "adjoint": {"value": npts / (adj_ms / 1e3), "unit": "force-pts/s",
                                           "avg_launch_ms": adj_ms, "cost_vs_forward": adj_ms / map_ms,
                                           "kernel": "mms_rate_grad_points (plan-specialised)",
                                           "compile_s": grad_compile_s,
                                           "outputs": "F + dF/d[det0..3, split, s0, B', z, v]",
                                           "generic_kernel": {"avg_launch_ms": gen_adj_ms,
                                                              "value": npts / (gen_adj_ms / 1e3),
                                                              "cost_vs_forward": gen_adj_ms / gen_map_ms,
                                                              "F_bit_identical": f_same,
                                                              "J_max_rel_diff": jrel}}}
        _progress("force map done")

    if not args.no_adjoint:
        result["ensemble_adjoint"] = adjoint_bench(mol, beams, z0, v0, B, dt, n_steps, stream)
        _progress("ensemble adjoint done")
    if not args.no_map:
        result["ensemble_size_scan"] = size_scan(mol, plan, B, dt, ks, stream)
        _progress("size scan done")

    if not args.no_traj3d:
        result["trajectory_models"] = traj3d_bench(mol, plan, B, dt, n_steps, stream, plan_generic)
        _progress("trajectory models done")
        result["rf_mot"] = rf_mot_bench(mol, z0, v0, dt, n_steps)
        _progress("rf mot done")

    if not args.no_lindblad:
        result["lindblad"] = lindblad_bench(mol, dev, stream, args.lindblad_steps)
        _progress("lindblad done")

    if not args.no_mcwf:
        result["mcwf"], mcwf_check = mcwf_bench(mol, dev, stream, args.mcwf_steps)
        _progress("mcwf done")

    checks = {}
    if not args.no_mcwf:
        checks["mcwf"] = mcwf_check
    if not args.no_floquet:
        result["floquet"], checks["floquet"] = floquet_bench(mol, dev, stream)
        _progress("floquet done")

    if not args.no_cpu:
        # CPU leg: the oracle timed as the baseline, and as the checker of
        # the GPU results above
        result["cpu_baseline"], ref = cpu_baseline(mol, beams, z0_h, v0_h, B, dt, n_steps)
        result["error_vs_oracle"] = ensemble_parity(plan, z0, v0, B, dt, n_steps, ref, stream)
        if not args.no_map:
            result["force_map"]["error_vs_oracle"] = map_parity(mol, beams, *map_out, B)
        if "floquet" in checks:
            result["floquet"]["force_err_vs_oracle"] = floquet_parity(mol, checks["floquet"])
        if "mcwf" in checks:
            result["mcwf"]["force_vs_lindblad_fp64"] = mcwf_master_parity(checks["mcwf"])
    print(json.dumps(result))


if __name__ == "__main__":
    main()
