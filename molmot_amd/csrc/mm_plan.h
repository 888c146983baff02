// Device-resident tables for the 1D rate-equation kernels.
//
// The plan plays the role of the reference's GPU data cache CUDAMolData
// (molmot_cuda/rate_equations_cuda.py:399-441): it is built once per
// (molecule, beam set) and lives in device memory; hot calls take it by
// pointer and allocate nothing.  Unlike the reference, the (ground, excited,
// beam) coupling terms with |d|^2 >= 1e-15 (the reference's skip rule,
// molmot/obe/rate_equations_jit.py:67-69) are pre-sorted by (ground,
// excited) pair so the kernel walks a compact, wave-uniform term list whose
// entries are read through the scalar cache.
#pragma once

#ifndef __HIPCC_RTC__
#include <stdint.h>
#endif

#define MM_MAX_G 36
#define MM_MAX_E 8
#define MM_MAX_S (MM_MAX_G + MM_MAX_E)
#define MM_MAX_B 32
#define MM_MAX_PAIRS (MM_MAX_G * MM_MAX_E)
#define MM_MAX_TERMS 2048
#define MM_MAX_COMP 8

// One coupling term: ground g -> excited e driven by beam b.
struct MMTerm {
    double freq;   // beam frequency (Hz)
    double kdir;   // beam k-vector z projection
    double C2;     // s0_b * |d_{g,e,q_b}|^2 * (Gamma/2)^2 / (2 pi)^2
    int beam;      // beam index
    int comp;      // frequency-component index of the beam
};

// Packed per-ground and per-pair records read with one 64-byte scalar load
// each (the packed path of rate_eval_fast; unit directions, n_e <= 4, every
// pair empty or of kind 1).
struct alignas(64) MMGroundRec {
    double E, zd;
    double br[4];      // BR[f][g], f < n_e
    unsigned mask;     // bit e: pair (g, e) has its four terms (= pair_mask[g])
    unsigned dirm;     // bit 4e + c: slot c of the component-ordered record of pair (g, e) is a -z beam
    double Gamma;      // the plan's Gamma, read with the record
};
struct alignas(64) MMPairRec {
    double f[4];       // beam frequencies: two +z, then two -z
    double c2[4];      // C2 of the same terms
};

struct MMRatePlanDev {
    int n_g, n_e, n_b, n_terms;
    int unit_dirs;     // every |kdir| == 1: Doppler shift is +-one shared value
    int force_dense;   // some excited state has no decay channel: reduced solve invalid
    int n_comp;
    int pad0;
    double Gamma, k, Geff, hbar;
    double gh, ghs;           // Gamma/2, (Gamma/2)^2
    double gq;                // (Gamma / (4 pi))^2: (Gamma/2)^2 in Hz^2
    double mk;                // -k
    double gh_geff;           // Gamma/2 * Gamma_eff
    double fk_gh_geff;        // hbar * k * Gamma/2 * Gamma_eff
    double inv_gh_geff;       // 1 / (Gamma/2 * Gamma_eff)
    double E[MM_MAX_S];       // energies (Hz)
    double zd[MM_MAX_S];      // Zeeman z diagonal
    double BR[MM_MAX_E][MM_MAX_G];
    double d2[MM_MAX_G][MM_MAX_E][3];
    double bfreq[MM_MAX_B], bkdir[MM_MAX_B], bs0[MM_MAX_B];
    int bpol[MM_MAX_B];
    int bcomp[MM_MAX_B];
    double split_sign[MM_MAX_COMP];
    int pair_start[MM_MAX_PAIRS + 1];
    int pair_kind[MM_MAX_PAIRS];  // 1: exactly four terms with kdir (+1, +1, -1, -1) in that order
    int packed;                    // the packed records below are valid
    unsigned pair_mask[MM_MAX_G];  // bit e: pair (g, e) has its four terms
    MMGroundRec grec[MM_MAX_G];
    MMPairRec prec[MM_MAX_G * 4];
    // The same four terms ordered by frequency component (slot c holds the
    // term of component c, its direction in MMGroundRec::dirm), valid when
    // every packed pair has one term of each component 0..3: the adjoint
    // then accumulates per-component sums without dispatching on the
    // component of each term.
    int comp_order;
    MMPairRec precc[MM_MAX_G * 4];
    MMTerm terms[MM_MAX_TERMS];
};
