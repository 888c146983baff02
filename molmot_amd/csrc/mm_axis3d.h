// Per-axis 3D model of the 1D rate-equation force
// (molmot/propagation/trajectories.py:78-156): each axis of a particle is one
// lane of a lane quad (x, y, z, idle).  Lane a evaluates F(v_a, z_eff) with
// z_eff = z for the z axis and -x/2, -y/2 for the transverse axes (the
// quadrupole field B = B'(-x/2, -y/2, z)); the escape radius |r| is reduced
// across the quad with two lane swaps, in the reference's summation order
// ((x^2 + y^2) + z^2).  The bodies of the two launch forms (one quad per
// particle; the work queue of mm_queue_dev.h) over an evaluator EV of the
// force (mm_traj.h's interface): the kernels with the tables in memory
// (traj3d.hip) and the plan-specialised ones (rate_spec.hip) share them.
// Free of host headers: run-time compiled kernels include it.
#pragma once

#include "mm_queue_dev.h"
#include "mm_traj.h"

// One quad per particle (k_traj_axis3d's contract).  RF: the force is the
// mean of P and of P2 at -Bgrad, the two switching phases of the RF MOT
// (molmot/mot/simulator.py:49-90).
template <class EV, int NG, int NE, bool UNIT, bool SAVE, int BS, bool RF>
__device__ __forceinline__ void axis3d_grid_body(const MMRatePlanDev* __restrict__ P,
                                                 const MMRatePlanDev* __restrict__ P2, long tid, long n,
                                                 const double* __restrict__ r0, const double* __restrict__ v0,
                                                 double Bgrad, double mass, double dt, int n_steps, double r_escape,
                                                 double* __restrict__ rt, double* __restrict__ vt,
                                                 double* __restrict__ rf, double* __restrict__ vf,
                                                 int* __restrict__ nact, double* lds) {
    const long i = tid >> 2;
    const int a = (int)(tid & 3);
    if (i >= n) return;  // whole quads leave together
    const double inv_mass = 1.0 / mass;
    const bool axis = a < 3;
    double r = axis ? r0[3 * i + a] : 0.0;
    double v = axis ? v0[3 * i + a] : 0.0;
    int na = n_steps;
    const long base = i * (long)n_steps * 3;
    if (SAVE && axis) {
        rt[base + a] = r;
        vt[base + a] = v;
    }
    for (int j = 1; j < n_steps; j++) {
        double f = 0.0;
        if (axis) {
            double rr;
            const double zeff = (a == 2) ? r : -r / 2.0;
            EV::eval(P, v, zeff, Bgrad, f, rr, lds, BS);
            if (RF) {
                double f2;
                rate_eval<NG, NE, UNIT>(P2, v, zeff, -Bgrad, f2, rr, nullptr, lds, BS);
                f = 0.5 * (f + f2);
            }
        }
        const double acc = mm_div_rn(f, mass, inv_mass);
        v = v + acc * dt;
        r = r + v * dt;
        if (SAVE && axis) {
            rt[base + (long)j * 3 + a] = r;
            vt[base + (long)j * 3 + a] = v;
        }
        const double sq = axis ? r * r : 0.0;
        const double s1 = sq + __shfl_xor(sq, 1);
        const double s2 = s1 + __shfl_xor(s1, 2);
        if (sqrt(s2) > r_escape) {
            na = j;
            if (SAVE && axis) {
                for (int k = j + 1; k < n_steps; k++) {
                    rt[base + (long)k * 3 + a] = r;
                    vt[base + (long)k * 3 + a] = v;
                }
            }
            break;
        }
    }
    if (axis) {
        rf[3 * i + a] = r;
        vf[3 * i + a] = v;
    }
    if (a == 0) nact[i] = na;
}

// Work-queue form (k_traj_axis3d_queue's contract; no saved trajectory, no
// RF): a group is the 16 particles of one wavefront (lane quads), a
// particle's r, v live in rf, vf and its n_actual in nact between chunks of
// steps; a quad continues while n_actual = n_steps.
template <class EV>
__device__ __forceinline__ void axis3d_queue_body(const MMRatePlanDev* __restrict__ P, long n,
                                                  const double* __restrict__ r0, const double* __restrict__ v0,
                                                  double Bgrad, double mass, double dt, int n_steps,
                                                  double r_escape, double* __restrict__ rf, double* __restrict__ vf,
                                                  int* __restrict__ nact, const MMQueue& Q, double* lds) {
    const double inv_mass = 1.0 / mass;
    const int a = (int)(threadIdx.x & 3);
    const bool axis = a < 3;
    unsigned g, ch;
    while (mm_queue_next(Q, g, ch)) {
        const long i = (long)g * 16 + (threadIdx.x >> 2);
        if (i < n) {  // whole quads together
            const int j0 = 1 + (int)ch * Q.chunk;
            const int j1 = min(n_steps, j0 + Q.chunk);
            double r, v;
            int na;
            if (ch == 0) {
                r = axis ? r0[3 * i + a] : 0.0;
                v = axis ? v0[3 * i + a] : 0.0;
                na = n_steps;
            } else {
                r = axis ? rf[3 * i + a] : 0.0;
                v = axis ? vf[3 * i + a] : 0.0;
                na = nact[i];
            }
            if (na == n_steps) {
                for (int j = j0; j < j1; j++) {
                    double f = 0.0;
                    if (axis) {
                        double rr;
                        const double zeff = (a == 2) ? r : -r / 2.0;
                        EV::eval(P, v, zeff, Bgrad, f, rr, lds, 64);
                    }
                    const double acc = mm_div_rn(f, mass, inv_mass);
                    v = v + acc * dt;
                    r = r + v * dt;
                    const double sq = axis ? r * r : 0.0;
                    const double s1 = sq + __shfl_xor(sq, 1);
                    const double s2 = s1 + __shfl_xor(s1, 2);
                    if (sqrt(s2) > r_escape) {
                        na = j;
                        break;
                    }
                }
            }
            if (axis) {
                rf[3 * i + a] = r;
                vf[3 * i + a] = v;
            }
            if (a == 0) nact[i] = na;
        }
        mm_queue_publish(Q, g, ch);
    }
}
