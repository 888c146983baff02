// Plan-specialised 1D trajectory kernels, compiled at run time with hipRTC.
//
// The reference compiles its GPU kernels at run time too (CuPy RawKernel
// source strings, molmot_cuda/trajectories_cuda.py:44-300), but with the
// molecule's tables passed as arrays.  Here a plan's tables become the
// constants of the code itself: for a packed plan (unit beam directions,
// every (ground, excited) pair empty or driven by two +z and two -z beams)
// the generator writes the rate-equation evaluation of rate_core.h's packed
// path as straight-line code over the pairs that are not empty, with every
// energy, Zeeman coefficient, beam frequency, coupling constant and
// branching ratio a literal.  What that removes from the generic kernel's
// step: the scalar loads of the pair and ground records and their waits,
// the per-pair branch on the pair mask (which serialised the pairs), the
// per-pair Doppler-shifted frequencies (each distinct freq +- dop is formed
// once per step), and every multiply-add by a branching ratio or rate that
// is zero.  The arithmetic of every remaining operation is the generic
// kernel's, in the same order (the terms left out are exact zeros: x + 0,
// fma(0, b, c) and fma(a, 0, c) return their other operand unchanged for the
// finite, non-negative values involved), so the specialised kernels give
// the generic kernels' trajectories bit for bit, and the detuning keeps the
// reference's rounding (rate_core.h).
//
// The step loop, the Philox kicks and the work queue are the generic
// kernels' own code (mm_traj.h, philox.h, mm_queue_dev.h), embedded in the
// library at build time and handed to hipRTC as headers.  A compiled module
// is cached per (device, generated source) and shared by every plan with the
// same tables; the last few that no plan references stay loaded.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/molmot_hip.h"
#include "mm_errors.h"
#include "mm_spec.h"
#include "rate_core.h"

// the device headers, embedded by molmot_amd/csrc/build.py
extern const char* const mm_rtc_header_names[];
extern const char* const mm_rtc_header_srcs[];
extern const int mm_rtc_n_headers;

// forward: mms_traj_queue_k0, _k1, mms_traj_grid_k0, _k1 (k: recoil kicks),
// mms_vjp_forward, mms_rate_map, mms_rate_points, mms_vsq_grid,
// mms_vsq_queue, mms_axis3d_grid, mms_axis3d_queue, mms_traj_stats_k0, _k1,
// mms_traj_save_k0, _k1 (saved trajectories); gradient:
// mms_tangent_queue, mms_tangent_grid, mms_vjp_segment, mms_rate_grad_points
struct MMSpecModule {
    int device, kind, refs;
    hipModule_t mod;
    hipFunction_t fn[15];
    std::pair<int, std::string> key;
};

static std::mutex mm_spec_mtx;
static std::map<std::pair<int, std::string>, MMSpecModule*> mm_spec_cache;
static std::vector<MMSpecModule*> mm_spec_idle;  // cached, referenced by no plan (oldest first)

// ---------------------------------------------------------------------------
// generator
// ---------------------------------------------------------------------------

// exact decimal-free literal of a double
static std::string lit(double x) {
    char b[64];
    snprintf(b, sizeof b, "(%a)", x);
    return b;
}

static void emit(std::string& o, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
static void emit(std::string& o, const char* fmt, ...) {
    char b[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(b, sizeof b, fmt, ap);
    va_end(ap);
    o += b;
}

// index of x among the distinct values in `vals` (bitwise), appending it
static int distinct_index(std::vector<double>& vals, double x) {
    for (size_t i = 0; i < vals.size(); i++)
        if (memcmp(&vals[i], &x, sizeof x) == 0) return (int)i;
    vals.push_back(x);
    return (int)vals.size() - 1;
}

// Why a plan cannot be specialised, or null.
static const char* spec_unsupported(const MMRatePlanDev& P) {
    if (!P.packed) return "the plan is not packed (unit beam directions, n_excited <= 4, every pair empty or "
                          "driven by two +z and two -z beams)";
    if (P.force_dense) return "an excited state has no decay channel (the reduced solve does not apply)";
    for (int g = 0; g < P.n_g; g++)
        if (P.pair_mask[g] == 0) return "a ground state is dark (every point takes the dense solve)";
    return nullptr;
}

// Where each pair's terms find their Doppler-shifted frequencies: the
// distinct values f + dop (+z beams) and f - dop (-z beams) are formed once
// per step; slot[(g*4 + e)*4 + j] indexes the term j of the packed record
// (two +z, then two -z), cslot the term of frequency component j of the
// component-ordered record (index into fp, or into fm when cdir is set).
struct SpecLayout {
    std::vector<double> fp, fm;
    std::vector<int> slot, cslot, cdir;
    std::vector<int> tslot;  // plans outside the packed form: term -> fp (kdir > 0) or fm index
};

static SpecLayout spec_layout(const MMRatePlanDev& P) {
    SpecLayout L;
    if (!P.packed) {
        L.tslot.assign(P.n_terms, -1);
        for (int it = 0; it < P.n_terms; it++)
            L.tslot[it] = P.terms[it].kdir > 0.0 ? distinct_index(L.fp, P.terms[it].freq)
                                                 : distinct_index(L.fm, P.terms[it].freq);
        return L;
    }
    L.slot.assign(P.n_g * 16, -1);
    L.cslot.assign(P.n_g * 16, -1);
    L.cdir.assign(P.n_g * 16, 0);
    for (int g = 0; g < P.n_g; g++)
        for (int e = 0; e < P.n_e; e++) {
            if (!(P.pair_mask[g] & (1u << e))) continue;
            const MMPairRec& Q = P.prec[g * 4 + e];
            for (int j = 0; j < 4; j++)
                L.slot[(g * 4 + e) * 4 + j] = j < 2 ? distinct_index(L.fp, Q.f[j]) : distinct_index(L.fm, Q.f[j]);
            if (!P.comp_order) continue;
            const MMPairRec& C = P.precc[g * 4 + e];
            for (int c = 0; c < 4; c++) {
                const bool minus = (P.grec[g].dirm >> (4 * e + c)) & 1u;
                L.cdir[(g * 4 + e) * 4 + c] = minus ? 1 : 0;
                L.cslot[(g * 4 + e) * 4 + c] = minus ? distinct_index(L.fm, C.f[c]) : distinct_index(L.fp, C.f[c]);
            }
        }
    return L;
}

// Bg, the Doppler shift, the excited energies and the shifted frequencies
// (the generic kernel's roundings, rate_core.h)
static void emit_prelude(std::string& o, const MMRatePlanDev& P, const SpecLayout& L) {
    const int NG = P.n_g;
    o += "  const double Bg = Bgrad * z * 1e4;\n";
    emit(o, "  const double dop = mm_div_rn(%s * v, MM_TWO_PI, MM_INV_TWO_PI);\n", lit(P.mk).c_str());
    for (int e = 0; e < P.n_e; e++) {
        const double E = P.E[NG + e], zd = P.zd[NG + e];
        if (zd == 0.0)
            emit(o, "  const double Ee%d = %s;\n", e, lit(E).c_str());
        else
            emit(o, "  const double Ee%d = %s + mm_div_rn(Bg * %s * %s, MM_TWO_PI, MM_INV_TWO_PI);\n", e,
                 lit(E).c_str(), lit(zd).c_str(), lit(P.Gamma).c_str());
    }
    for (size_t i = 0; i < L.fp.size(); i++) emit(o, "  const double fp%zu = %s + dop;\n", i, lit(L.fp[i]).c_str());
    for (size_t i = 0; i < L.fm.size(); i++) emit(o, "  const double fm%zu = %s - dop;\n", i, lit(L.fm[i]).c_str());
    o += "  double K[MMS_NE][MMS_NE];\n";
    o += "#pragma unroll\n  for (int e = 0; e < MMS_NE; e++)\n#pragma unroll\n"
         "    for (int f = 0; f < MMS_NE; f++) K[e][f] = 0.0;\n";
    o += "  bool ok = true;\n";
}

static std::string ground_energy(const MMRatePlanDev& P, int g) {
    const double E = P.E[g], zd = P.zd[g];
    if (zd == 0.0) return lit(E);
    char b[256];
    snprintf(b, sizeof b, "%s + mm_div_rn(Bg * %s * %s, MM_TWO_PI, MM_INV_TWO_PI)", lit(E).c_str(), lit(zd).c_str(),
             lit(P.Gamma).c_str());
    return b;
}

static std::vector<int> coupled(const MMRatePlanDev& P, int g) {
    std::vector<int> es;
    for (int e = 0; e < P.n_e; e++) {
        const int pr = g * P.n_e + e;
        if (P.packed ? (P.pair_mask[g] & (1u << e)) != 0 : P.pair_start[pr + 1] > P.pair_start[pr]) es.push_back(e);
    }
    return es;
}

// mm_pair_sums of rate_core.h (the general path, unit directions) for pair
// (g, e) of a plan outside the packed form: a (+z, +z, -z, -z) pair by one
// simultaneous inversion, any other term list in groups of four and then
// one term at a time, with the generic kernel's roundings and order
static void emit_pair_general(std::string& o, const MMRatePlanDev& P, const SpecLayout& L, int g, int e,
                              const char* eg, const std::string& sn, const std::string& tn) {
    const int pr = g * P.n_e + e;
    const int t0 = P.pair_start[pr], t1 = P.pair_start[pr + 1];
    const std::string gq = lit(P.gq);
    auto fd = [&](int it) {
        return std::string(P.terms[it].kdir > 0.0 ? "fp" : "fm") + std::to_string(L.tslot[it]);
    };
    emit(o, "    {\n      const double om = Ee%d - %s;\n", e, eg);
    if (P.pair_kind[pr] == 1) {
        emit(o, "      const double da = %s - om, db = %s - om, dc = %s - om, dd = %s - om;\n", fd(t0).c_str(),
             fd(t0 + 1).c_str(), fd(t0 + 2).c_str(), fd(t0 + 3).c_str());
        emit(o, "      const double xa = fma(da, da, %s), xb = fma(db, db, %s);\n", gq.c_str(), gq.c_str());
        emit(o, "      const double xc = fma(dc, dc, %s), xd = fma(dd, dd, %s);\n", gq.c_str(), gq.c_str());
        o += "      const double pab = xa * xb, pcd = xc * xd;\n";
        o += "      const double r = mm_rcp1(pab * pcd);\n";
        o += "      const double rab = r * pcd, rcd = r * pab;\n";
        emit(o, "      const double wa = %s * (xb * rab), wb = %s * (xa * rab);\n", lit(P.terms[t0].C2).c_str(),
             lit(P.terms[t0 + 1].C2).c_str());
        emit(o, "      const double wc = %s * (xd * rcd), wd = %s * (xc * rcd);\n", lit(P.terms[t0 + 2].C2).c_str(),
             lit(P.terms[t0 + 3].C2).c_str());
        o += "      const double sp = wa + wb, sm = wc + wd;\n";
        emit(o, "      %s = sp + sm;\n      %s = sp - sm;\n    }\n", sn.c_str(), tn.c_str());
        return;
    }
    o += "      double s = 0.0, t = 0.0;\n";
    int it = t0;
    static const char* nm[4] = {"a", "b", "c", "d"};
    for (; it + 4 <= t1; it += 4) {
        o += "      {\n";
        for (int j = 0; j < 4; j++) {
            emit(o, "        const double d%s = %s - om;\n", nm[j], fd(it + j).c_str());
            emit(o, "        const double w%s = %s * mm_rcp1(fma(d%s, d%s, %s));\n", nm[j],
                 lit(P.terms[it + j].C2).c_str(), nm[j], nm[j], gq.c_str());
        }
        o += "        s += (wa + wb) + (wc + wd);\n";
        emit(o, "        t += fma(%s, wa, %s * wb) + fma(%s, wc, %s * wd);\n      }\n", lit(P.terms[it].kdir).c_str(),
             lit(P.terms[it + 1].kdir).c_str(), lit(P.terms[it + 2].kdir).c_str(), lit(P.terms[it + 3].kdir).c_str());
    }
    for (; it < t1; ++it) {
        emit(o, "      {\n        const double d = %s - om;\n", fd(it).c_str());
        emit(o, "        const double w = %s * mm_rcp1(fma(d, d, %s));\n", lit(P.terms[it].C2).c_str(), gq.c_str());
        emit(o, "        s += w;\n        t = fma(%s, w, t);\n      }\n", lit(P.terms[it].kdir).c_str());
    }
    emit(o, "      %s = s;\n      %s = t;\n    }\n", sn.c_str(), tn.c_str());
}

// mm_pair4 of rate_core.h for pair (g, e): S into `sn`, T into `tn`
static void emit_pair_forward(std::string& o, const MMRatePlanDev& P, const SpecLayout& L, int g, int e,
                              const char* eg, const std::string& sn, const std::string& tn) {
    const MMPairRec& Q = P.prec[g * 4 + e];
    const int* sl = &L.slot[(g * 4 + e) * 4];
    const std::string gq = lit(P.gq);
    emit(o, "    {\n      const double om = Ee%d - %s;\n", e, eg);
    emit(o, "      const double da = fp%d - om, db = fp%d - om, dc = fm%d - om, dd = fm%d - om;\n", sl[0], sl[1],
         sl[2], sl[3]);
    emit(o, "      const double xa = fma(da, da, %s), xb = fma(db, db, %s);\n", gq.c_str(), gq.c_str());
    emit(o, "      const double xc = fma(dc, dc, %s), xd = fma(dd, dd, %s);\n", gq.c_str(), gq.c_str());
    o += "      const double pab = xa * xb, pcd = xc * xd;\n";
    o += "      const double r = mm_rcp1(pab * pcd);\n";
    emit(o, "      const double np = fma(%s, xb, %s * xa) * pcd;\n", lit(Q.c2[0]).c_str(), lit(Q.c2[1]).c_str());
    emit(o, "      const double nm = fma(%s, xd, %s * xc) * pab;\n", lit(Q.c2[2]).c_str(), lit(Q.c2[3]).c_str());
    emit(o, "      %s = (np + nm) * r;\n      %s = (np - nm) * r;\n    }\n", sn.c_str(), tn.c_str());
}

// left-to-right sum of names[] (the generic kernel's order)
static std::string sum_of(const std::vector<std::string>& names) {
    std::string s = names[0];
    for (size_t k = 1; k < names.size(); k++) s = "(" + s + " + " + names[k] + ")";
    return s;
}

// K[e][f] += (S_ge / S_g) BR[f][g] for the coupled pairs of g and the
// non-zero branching ratios
static void emit_k_update(std::string& o, const MMRatePlanDev& P, int g, const std::vector<std::string>& sn) {
    const std::vector<int> es = coupled(P, g);
    for (size_t k = 0; k < es.size(); k++) {
        const int e = es[k];
        bool any = false;
        for (int f = 0; f < P.n_e; f++) any = any || (f != e && P.BR[f][g] != 0.0);
        if (!any) continue;
        emit(o, "    {\n      const double a = %s * rS;\n", sn[k].c_str());
        for (int f = 0; f < P.n_e; f++)
            if (f != e && P.BR[f][g] != 0.0)
                emit(o, "      K[%d][%d] = fma(a, %s, K[%d][%d]);\n", e, f, lit(P.BR[f][g]).c_str(), e, f);
        o += "    }\n";
    }
}

// u = sum_f y[f] BR[f][g] (non-zero ratios, generic order) into `name`
static void emit_u(std::string& o, const MMRatePlanDev& P, int g, const char* yname, const char* name) {
    emit(o, "    double %s = 0.0;\n", name);
    for (int f = 0; f < P.n_e; f++)
        if (P.BR[f][g] != 0.0) emit(o, "    %s = fma(%s[%d], %s, %s);\n", name, yname, f, lit(P.BR[f][g]).c_str(), name);
}

// The forward evaluation (rate_core.h's packed path; same F and R bits);
// with_pop: mms_fast_pop, which also writes the NG ground populations.
static void emit_fast(std::string& o, const MMRatePlanDev& P, const SpecLayout& L, bool with_pop = false,
                      const char* name = "mms_fast") {
    const int NG = P.n_g;
    if (with_pop)
        o += "__device__ __forceinline__ bool mms_fast_pop(double v, double z, double Bgrad, double& F, double& R,\n"
             "                                             double* __restrict__ pop) {\n";
    else
        emit(o, "__device__ __forceinline__ bool %s(double v, double z, double Bgrad, double& F, double& R) {\n",
             name);
    emit_prelude(o, P, L);
    for (int g = 0; g < NG; g++) emit(o, "  double D%d, G%d;\n", g, g);
    for (int g = 0; g < NG; g++) {
        emit(o, "  { // ground state %d\n", g);
        emit(o, "    const double Eg = %s;\n", ground_energy(P, g).c_str());
        const std::vector<int> es = coupled(P, g);
        std::vector<std::string> sn, tn;
        for (int e : es) {
            sn.push_back("ps" + std::to_string(e));
            tn.push_back("pt" + std::to_string(e));
            emit(o, "    double ps%d, pt%d;\n", e, e);
            if (P.packed)
                emit_pair_forward(o, P, L, g, e, "Eg", sn.back(), tn.back());
            else
                emit_pair_general(o, P, L, g, e, "Eg", sn.back(), tn.back());
        }
        // Sg, Tg: the generic kernel's sums over all four pairs, whose empty
        // pairs add exact zeros
        emit(o, "    const double Sg = %s, Tg = %s;\n", sum_of(sn).c_str(), sum_of(tn).c_str());
        o += "    ok = ok && (Sg > 0.0);\n";
        o += "    const double s1 = 1.0 + Sg;\n    const double rr = mm_rcp(Sg * s1);\n";
        o += "    const double rS = rr * s1, sat = rr * Sg;\n";
        emit_k_update(o, P, g, sn);
        emit(o, "    D%d = (rS * s1) * %s;\n", g, lit(P.inv_gh_geff).c_str());
        emit(o, "    G%d = %s * sat * Tg;\n  }\n", g, lit(P.fk_gh_geff).c_str());
    }
    o += "  double y[MMS_NE];\n  mm_gth_stationary<MMS_NE>(K, ok, y);\n";
    o += "  double Z = 0.0, Fs = 0.0, Rs = 0.0;\n";
    for (int g = 0; g < NG; g++) {
        if (with_pop) emit(o, "  double q%d;\n", g);
        o += "  {\n";
        emit_u(o, P, g, "y", "u");
        emit(o, "    const double q = u * D%d;\n    Z += q;\n    Fs = fma(q, G%d, Fs);\n    Rs += u;\n", g, g);
        if (with_pop) emit(o, "    q%d = q;\n", g);
        o += "  }\n";
    }
    o += "  const double rZ = mm_rcp(Z);\n  F = Fs * rZ;\n  R = Rs * rZ;\n";
    if (with_pop)
        for (int g = 0; g < NG; g++) emit(o, "  pop[%d] = q%d * rZ;\n", g, g);
    o += "  return ok;\n}\n\n";
}

// F and dF/d[4 detunings, split, s0 scale, B', z, v] (rate_adjoint.h's
// reverse pass, restructured for straight-line code).  The forward pass is
// rate_adjoint.h's (its F bits); the reverse pass forms the complete adjoint
// of S_g before the pairs, from the forward's pair sums, so that every term's
// weight W = (L_g - lambda'_e u_g)/S_g + Sbar_g +- Tbar_g is known when the
// term is re-evaluated: dF/dd = -2 D W with D = w d / (d^2 + gq), and the
// sums over the terms reduce to one multiply-add into the term's component
// and two direction sums per pair.
static void emit_grad(std::string& o, const MMRatePlanDev& P, const SpecLayout& L) {
    const int NG = P.n_g, NE = P.n_e;
    const double zfac = P.Gamma / MM_TWO_PI;
    const std::string c1 = lit(P.gh_geff), c2 = lit(P.fk_gh_geff), igg = lit(P.inv_gh_geff), gq = lit(P.gq);
    o += "__device__ __forceinline__ bool mms_grad(double v, double z, double Bgrad, double& F,\n"
         "                                         double (&J)[MM_ADJ_NJ]) {\n";
    emit_prelude(o, P, L);
    // ---- forward: per ground rS, sat, Bsat = sat T_g, kept in a per-lane LDS
    // slab (64-lane blocks) so that registers stay under half a SIMD's file
    o += "  __shared__ double mms_slab[3 * MMS_NG * 64];\n  double* const sl = mms_slab + threadIdx.x;\n";
    for (int g = 0; g < NG; g++) {
        emit(o, "  { // ground state %d\n", g);
        emit(o, "    const double Eg = %s;\n", ground_energy(P, g).c_str());
        const std::vector<int> es = coupled(P, g);
        std::vector<std::string> sn, tn;
        for (int e : es) {
            sn.push_back("ps" + std::to_string(e));
            tn.push_back("pt" + std::to_string(e));
            emit(o, "    double ps%d, pt%d;\n", e, e);
            emit_pair_forward(o, P, L, g, e, "Eg", sn.back(), tn.back());
        }
        emit(o, "    const double Sg = %s, Tg = %s;\n", sum_of(sn).c_str(), sum_of(tn).c_str());
        o += "    ok = ok && (Sg > 0.0);\n";
        o += "    const double s1 = 1.0 + Sg;\n    const double rr = mm_rcp(Sg * s1);\n";
        o += "    const double rS = rr * s1;\n";
        emit_k_update(o, P, g, sn);
        o += "    const double sat = rr * Sg;\n";
        emit(o, "    sl[%d] = rS;\n    sl[%d] = sat;\n    sl[%d] = sat * Tg;\n  }\n", (0 * NG + g) * 64, (1 * NG + g) * 64,
             (2 * NG + g) * 64);
    }
    // generator of the excited chain (kept for the multiplier solve), GTH
    o += R"(  double G[MMS_NE][MMS_NE];
#pragma unroll
  for (int e = 0; e < MMS_NE; e++) {
    double out = 0.0;
#pragma unroll
    for (int f = 0; f < MMS_NE; f++) {
      G[e][f] = (f != e) ? K[e][f] : 0.0;
      if (f != e) out += K[f][e];
    }
    G[e][e] = -out;
  }
  double y[MMS_NE];
  mm_gth_stationary<MMS_NE>(K, ok, y);
  double Z = 0.0, Fs = 0.0;
)";
    for (int g = 0; g < NG; g++) {
        o += "  {\n";
        emit_u(o, P, g, "y", "u");
        emit(o, "    const double q = u * ((1.0 + sl[%d]) * %s);\n    Z += q;\n    Fs = fma(q, %s * sl[%d], Fs);\n  }\n",
             (0 * NG + g) * 64, igg.c_str(), c2.c_str(), (2 * NG + g) * 64);
    }
    o += "  const double rZ = mm_rcp(Z);\n  F = Fs * rZ;\n";
    // ---- ybar and the multipliers lambda (rate_adjoint.h)
    o += "  double yb[MMS_NE];\n#pragma unroll\n  for (int f = 0; f < MMS_NE; f++) yb[f] = 0.0;\n";
    for (int g = 0; g < NG; g++) {
        emit(o, "  {\n    const double ub = (%s * sl[%d] - F) * rZ * ((1.0 + sl[%d]) * %s);\n", c2.c_str(),
             (2 * NG + g) * 64, (0 * NG + g) * 64, igg.c_str());
        for (int f = 0; f < NE; f++)
            if (P.BR[f][g] != 0.0) emit(o, "    yb[%d] = fma(ub, %s, yb[%d]);\n", f, lit(P.BR[f][g]).c_str(), f);
        o += "  }\n";
    }
    o += R"(  double A[MMS_NE][MMS_NE], lam[MMS_NE], rdg[MMS_NE];
#pragma unroll
  for (int i = 0; i < MMS_NE; i++) {
#pragma unroll
    for (int j = 0; j < MMS_NE; j++) A[i][j] = (j == 0) ? (i == 0 ? 1.0 : 0.0) : G[j][i];
    lam[i] = yb[i];
  }
  rdg[0] = 1.0;
#pragma unroll
  for (int c = 1; c < MMS_NE; c++) {
    rdg[c] = mm_rcp1(A[c][c]);
#pragma unroll
    for (int r = c + 1; r < MMS_NE; r++) {
      const double fct = A[r][c] * rdg[c];
#pragma unroll
      for (int j = c + 1; j < MMS_NE; j++) A[r][j] = fma(-fct, A[c][j], A[r][j]);
      lam[r] = fma(-fct, lam[c], lam[r]);
    }
  }
#pragma unroll
  for (int i = MMS_NE - 1; i >= 0; i--) {
    double s = lam[i];
#pragma unroll
    for (int j = i + 1; j < MMS_NE; j++) s = fma(-A[i][j], lam[j], s);
    lam[i] = s * rdg[i];
  }
  double ly[MMS_NE];
#pragma unroll
  for (int e = 0; e < MMS_NE; e++) ly[e] = e > 0 ? lam[e] * y[e] : 0.0;
  double dc0 = 0.0, dc1 = 0.0, dc2 = 0.0, dc3 = 0.0, dopa = 0.0, Bga = 0.0, lamb = 0.0;
)";
    // ---- reverse, per ground state.  The terms are re-evaluated from
    // copies of the shifted frequencies that the compiler cannot identify
    // with the forward pass's (an empty asm statement): otherwise it keeps
    // every forward term value alive through the multiplier solve, which
    // needs more registers than a SIMD has per lane
    for (size_t i = 0; i < L.fp.size(); i++)
        emit(o, "  double rfp%zu = fp%zu;\n  asm volatile(\"\" : \"+v\"(rfp%zu));\n", i, i, i);
    for (size_t i = 0; i < L.fm.size(); i++)
        emit(o, "  double rfm%zu = fm%zu;\n  asm volatile(\"\" : \"+v\"(rfm%zu));\n", i, i, i);
    // likewise the stationary vector, so that the per-ground products of
    // the forward pass (u_g, q_g, ...) are not kept for the reverse pass
    o += "  double ry[MMS_NE];\n#pragma unroll\n  for (int f = 0; f < MMS_NE; f++) {\n"
         "    ry[f] = y[f];\n    asm volatile(\"\" : \"+v\"(ry[f]));\n  }\n";
    for (int g = 0; g < NG; g++) {
        const std::vector<int> es = coupled(P, g);
        emit(o, "  { // ground state %d\n", g);
        emit(o, "    const double rSg = sl[%d], satg = sl[%d], Bsg = sl[%d];\n", (0 * NG + g) * 64, (1 * NG + g) * 64,
             (2 * NG + g) * 64);
        emit_u(o, P, g, "ry", "u");
        emit_u(o, P, g, "ly", "Lg");
        emit(o, "    const double rD = (1.0 + rSg) * %s;\n", igg.c_str());
        o += "    const double q = u * rD;\n";
        emit(o, "    const double qb = (%s * Bsg - F) * rZ;\n", c2.c_str());
        o += "    const double Db = -qb * q * rD;\n    const double Fgb = q * rZ;\n";
        emit(o, "    const double Tb = Fgb * %s * satg;\n", c2.c_str());
        emit(o, "    double Sb = Db * %s * satg * satg - Fgb * %s * Bsg * satg;\n", c1.c_str(), c2.c_str());
        emit(o, "    const double Eg = %s;\n", ground_energy(P, g).c_str());
        // phase A: each coupled pair's terms re-evaluated (component order)
        // and their sum s_e, which completes Sbar_g
        for (int e : es) {
            const int* cs = &L.cslot[(g * 4 + e) * 4];
            const int* cd = &L.cdir[(g * 4 + e) * 4];
            const MMPairRec& C = P.precc[g * 4 + e];
            emit(o, "    const double om%d = Ee%d - Eg;\n", e, e);
            for (int c = 0; c < 4; c++)
                emit(o, "    const double d%d_%d = %s%d - om%d;\n    const double x%d_%d = fma(d%d_%d, d%d_%d, %s);\n", e, c,
                     cd[c] ? "rfm" : "rfp", cs[c], e, e, c, e, c, e, c, gq.c_str());
            emit(o, "    const double p01_%d = x%d_0 * x%d_1, p23_%d = x%d_2 * x%d_3;\n", e, e, e, e, e, e);
            emit(o, "    const double rr%d = mm_rcp1(p01_%d * p23_%d);\n", e, e, e);
            emit(o, "    const double ra%d = rr%d * p23_%d, rb%d = rr%d * p01_%d;\n", e, e, e, e, e, e);
            emit(o, "    const double r%d_0 = x%d_1 * ra%d, r%d_1 = x%d_0 * ra%d, r%d_2 = x%d_3 * rb%d, r%d_3 = x%d_2 * rb%d;\n",
                 e, e, e, e, e, e, e, e, e, e, e, e);
            for (int c = 0; c < 4; c++) {
                emit(o, "    const double w%d_%d = %s * r%d_%d;\n", e, c, lit(C.c2[c]).c_str(), e, c);
                emit(o, "    const double D%d_%d = (w%d_%d * r%d_%d) * d%d_%d;\n", e, c, e, c, e, c, e, c);
            }
            emit(o, "    const double s%d = ((w%d_0 + w%d_1) + w%d_2) + w%d_3;\n", e, e, e, e, e);
            if (e == 0)
                emit(o, "    const double vb%d = Lg * rSg;\n", e);
            else
                emit(o, "    const double vb%d = fma(-lam[%d], u, Lg) * rSg;\n", e, e);
            emit(o, "    Sb = fma(-vb%d * rSg, s%d, Sb);\n", e, e);
        }
        o += "    const double SbP = Sb + Tb, SbM = Sb - Tb;\n";
        // phase B: dF/dd = -2 D W per term into its component, the B' and
        // Doppler sums and the s0-scale sum
        for (int e : es) {
            const int* cd = &L.cdir[(g * 4 + e) * 4];
            const double dz = P.zd[NG + e] * zfac - P.zd[g] * zfac;
            emit(o, "    {\n      const double Wp = vb%d + SbP, Wm = vb%d + SbM;\n", e, e);
            std::vector<std::string> dp_, dm_, wp_, wm_;
            for (int c = 0; c < 4; c++) {
                emit(o, "      dc%d = fma(D%d_%d, %s, dc%d);\n", c, e, c, cd[c] ? "Wm" : "Wp", c);
                (cd[c] ? dm_ : dp_).push_back("D" + std::to_string(e) + "_" + std::to_string(c));
                (cd[c] ? wm_ : wp_).push_back("w" + std::to_string(e) + "_" + std::to_string(c));
            }
            const std::string DP = dp_.empty() ? "0.0" : sum_of(dp_), DM = dm_.empty() ? "0.0" : sum_of(dm_);
            const std::string WP = wp_.empty() ? "0.0" : sum_of(wp_), WM = wm_.empty() ? "0.0" : sum_of(wm_);
            emit(o, "      const double DP = %s, DM = %s;\n", DP.c_str(), DM.c_str());
            o += "      const double aP = DP * Wp, aM = DM * Wm;\n";
            emit(o, "      Bga = fma(aP + aM, %s, Bga);\n", lit(dz).c_str());
            o += "      dopa += aP - aM;\n";
            emit(o, "      lamb = fma(%s, Wp, fma(%s, Wm, lamb));\n    }\n", WP.c_str(), WM.c_str());
        }
        o += "  }\n";
    }
    // ---- assembly (rate_adjoint.h's J layout and units)
    o += "  J[0] = -2.0 * dc0;\n  J[1] = -2.0 * dc1;\n  J[2] = -2.0 * dc2;\n  J[3] = -2.0 * dc3;\n";
    o += "  double split = 0.0;\n";
    for (int c = 0; c < 4; c++)
        emit(o, "  split = fma(%s, J[%d], split);\n", lit(c < P.n_comp ? P.split_sign[c] : 0.0).c_str(), c);
    o += "  J[MM_ADJ_J_SPLIT] = split;\n  J[MM_ADJ_J_S0] = lamb;\n";
    o += "  J[MM_ADJ_J_B] = (2.0 * Bga) * z * 1e4;\n  J[MM_ADJ_J_Z] = (2.0 * Bga) * Bgrad * 1e4;\n";
    emit(o, "  J[MM_ADJ_J_V] = (-2.0 * dopa) * %s / MM_TWO_PI;\n  return ok;\n}\n\n", lit(P.mk).c_str());
}

// The gradient kernels need the component-ordered records (one term of
// each of four frequency components per coupled pair).
static bool spec_has_grad(const MMRatePlanDev& P) { return P.comp_order && P.n_comp <= 4; }

// Source of the specialised kernels of one kind: MM_SPEC_FORWARD, the
// trajectory kernels (queue and one thread per particle, with and without
// recoil kicks); MM_SPEC_GRADIENT, the forward-sensitivity kernels.
static std::string spec_source(const MMRatePlanDev& P, int kind) {
    std::string o;
    o += kind == MM_SPEC_FORWARD ? "// generated: plan-specialised 1D trajectory kernels\n"
                                 : "// generated: plan-specialised 1D forward-sensitivity kernels\n";
    // hipRTC declares the fixed-width integer types in its own namespace
    o += "typedef __hip_internal::uint32_t uint32_t;\ntypedef __hip_internal::uint64_t uint64_t;\n";
    o += kind == MM_SPEC_FORWARD ? "#include \"mm_traj.h\"\n#include \"mm_vjp.h\"\n#include \"mm_vsq.h\"\n#include \"mm_axis3d.h\"\n"
                                 : "#include \"mm_vjp.h\"\n";
    emit(o, "#define MMS_NG %d\n#define MMS_NE %d\n", P.n_g, P.n_e);
    const SpecLayout L = spec_layout(P);
    if (kind == MM_SPEC_FORWARD) {
        emit_fast(o, P, L);
        emit_fast(o, P, L, true);
        o += R"(struct MMEvalSpec {
  __device__ __forceinline__ static void eval(const MMRatePlanDev* __restrict__ P, double v, double z, double Bgrad,
                                              double& F, double& R, double*, int) {
    if (!mms_fast(v, z, Bgrad, F, R)) {
      double Fd, Rd;
      rate_eval_dense<MMS_NG, MMS_NE>(P, v, z, Bgrad, &Fd, &Rd, nullptr);
      F = Fd;
      R = Rd;
    }
  }
};

#define MMS_QUEUE(NAME, KICKS)                                                                                      \
  extern "C" __global__ __launch_bounds__(64) void NAME(                                                           \
      const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,    \
      double Bgrad, double mass, double dt, int n_steps, double z_escape, double* __restrict__ zf,                  \
      double* __restrict__ vf, int* __restrict__ nact, double kick_scale, uint32_t k0, uint32_t k1, uint32_t sid, \
      MMQueue Q) {                                                                                                  \
    traj_queue_body<MMEvalSpec, MMS_NG, MMS_NE, true, KICKS>(P, n, z0, v0, Bgrad, mass, dt, n_steps, z_escape, zf, \
                                                             vf, nact, kick_scale, k0, k1, sid, Q, nullptr);        \
  }
#define MMS_GRID(NAME, KICKS)                                                                                       \
  extern "C" __global__ __launch_bounds__(64) void NAME(                                                           \
      const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,    \
      double Bgrad, double mass, double dt, int n_steps, double z_escape, double* __restrict__ zf,                  \
      double* __restrict__ vf, int* __restrict__ nact, double kick_scale, uint32_t k0, uint32_t k1, uint32_t sid) { \
    const long i = (long)blockIdx.x * 64 + threadIdx.x;                                                            \
    traj_grid_body<MMEvalSpec, MMS_NG, MMS_NE, true, false, KICKS, 64, false>(                                     \
        P, i, n, z0, v0, Bgrad, mass, dt, n_steps, z_escape, nullptr, nullptr, zf, vf, nact, kick_scale, k0, k1,   \
        sid, nullptr, nullptr);                                                                                     \
  }
#define MMS_GRID_SAVE(NAME, KICKS)                                                                                  \
  extern "C" __global__ __launch_bounds__(64) void NAME(                                                           \
      const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,    \
      double Bgrad, double mass, double dt, int n_steps, double z_escape, double* __restrict__ zt,                  \
      double* __restrict__ vt, double* __restrict__ zf, double* __restrict__ vf, int* __restrict__ nact,           \
      double kick_scale, uint32_t k0, uint32_t k1, uint32_t sid) {                                                  \
    const long i = (long)blockIdx.x * 64 + threadIdx.x;                                                            \
    traj_grid_body<MMEvalSpec, MMS_NG, MMS_NE, true, true, KICKS, 64, false>(                                      \
        P, i, n, z0, v0, Bgrad, mass, dt, n_steps, z_escape, zt, vt, zf, vf, nact, kick_scale, k0, k1, sid,        \
        nullptr, nullptr);                                                                                          \
  }
MMS_QUEUE(mms_traj_queue_k0, false)
MMS_QUEUE(mms_traj_queue_k1, true)
MMS_GRID(mms_traj_grid_k0, false)
MMS_GRID(mms_traj_grid_k1, true)
MMS_GRID_SAVE(mms_traj_save_k0, false)
MMS_GRID_SAVE(mms_traj_save_k1, true)

extern "C" __global__ __launch_bounds__(64) void mms_vjp_forward(const MMRatePlanDev* __restrict__ P, VjpArgs A) {
  vjp_forward_body<MMEvalSpec>(P, A, (long)blockIdx.x * 64 + threadIdx.x);
}

// F (and R, populations) at one point, with the dense fallback
__device__ __forceinline__ void mms_point(const MMRatePlanDev* __restrict__ P, double v, double z, double Bgrad,
                                          double& F, double& R, double* __restrict__ pop) {
  const bool ok = pop ? mms_fast_pop(v, z, Bgrad, F, R, pop) : mms_fast(v, z, Bgrad, F, R);
  if (!ok) {
    double Fd, Rd;
    rate_eval_dense<MMS_NG, MMS_NE>(P, v, z, Bgrad, &Fd, &Rd, pop);
    F = Fd;
    R = Rd;
  }
}

// the (v, z) map (k_rate_map's contract: index iv * n_z + iz)
extern "C" __global__ __launch_bounds__(64) void mms_rate_map(
    const MMRatePlanDev* __restrict__ P, const double* __restrict__ v, const double* __restrict__ z, int nz, long n,
    double Bgrad, double* __restrict__ F, double* __restrict__ pop, double* __restrict__ R) {
  const long idx = (long)blockIdx.x * 64 + threadIdx.x;
  if (idx >= n) return;
  const long iv = idx / nz;
  const long iz = idx - iv * nz;
  double f, r;
  mms_point(P, v[iv], z[iz], Bgrad, f, r, pop ? pop + idx * MMS_NG : nullptr);
  F[idx] = f;
  if (R) R[idx] = r;
}

// scattered points (k_rate_points's contract)
extern "C" __global__ __launch_bounds__(64) void mms_rate_points(
    const MMRatePlanDev* __restrict__ P, const double* __restrict__ v, const double* __restrict__ z, long n,
    double Bgrad, double* __restrict__ F, double* __restrict__ pop, double* __restrict__ R) {
  const long idx = (long)blockIdx.x * 64 + threadIdx.x;
  if (idx >= n) return;
  double f, r;
  mms_point(P, v[idx], z[idx], Bgrad, f, r, pop ? pop + idx * MMS_NG : nullptr);
  F[idx] = f;
  if (R) R[idx] = r;
}

// Monte-Carlo temperature statistics (k_traj_vsq's and k_traj_vsq_queue's contracts)
extern "C" __global__ __launch_bounds__(64) void mms_vsq_grid(
    const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,
    double Bgrad, double mass, double dt, int n_steps, int i_eq, double z_escape, double* __restrict__ vsq,
    long long* __restrict__ cnt) {
  vsq_grid_body<MMEvalSpec, MMS_NG, MMS_NE, true, 64, false>(P, nullptr, (long)blockIdx.x * 64 + threadIdx.x, n, z0,
                                                             v0, Bgrad, mass, dt, n_steps, i_eq, z_escape, vsq, cnt,
                                                             nullptr);
}
extern "C" __global__ __launch_bounds__(64) void mms_vsq_queue(
    const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,
    double Bgrad, double mass, double dt, int n_steps, int i_eq, double z_escape, double* __restrict__ vsq,
    long long* __restrict__ cnt, double* __restrict__ zs, double* __restrict__ vs, MMQueue Q) {
  vsq_queue_body<MMEvalSpec, MMS_NG, MMS_NE, true>(P, n, z0, v0, Bgrad, mass, dt, n_steps, i_eq, z_escape, vsq, cnt,
                                                   zs, vs, Q, nullptr);
}

// ensemble statistics at sample steps (k_traj_stats's contract; k: recoil kicks)
#define MMS_STATS(NAME, KICKS)                                                                                      \
  extern "C" __global__ __launch_bounds__(64) void NAME(                                                           \
      const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,    \
      double Bgrad, double mass, double dt, int n_steps, double z_escape, double r_alive,                          \
      const int* __restrict__ samp, int n_samp, double* __restrict__ part, double kick_scale, uint32_t k0,         \
      uint32_t k1, uint32_t sid) {                                                                                  \
    traj_stats_body<MMEvalSpec, MMS_NG, MMS_NE, true, KICKS>(P, n, z0, v0, Bgrad, mass, dt, n_steps, z_escape,     \
                                                             r_alive, samp, n_samp, part, kick_scale, k0, k1, sid, \
                                                             nullptr);                                              \
  }
MMS_STATS(mms_traj_stats_k0, false)
MMS_STATS(mms_traj_stats_k1, true)

// per-axis 3D model (k_traj_axis3d's finals-only and k_traj_axis3d_queue's contracts)
extern "C" __global__ __launch_bounds__(64) void mms_axis3d_grid(
    const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ r0, const double* __restrict__ v0,
    double Bgrad, double mass, double dt, int n_steps, double r_escape, double* __restrict__ rf,
    double* __restrict__ vf, int* __restrict__ nact) {
  axis3d_grid_body<MMEvalSpec, MMS_NG, MMS_NE, true, false, 64, false>(
      P, nullptr, (long)blockIdx.x * 64 + threadIdx.x, n, r0, v0, Bgrad, mass, dt, n_steps, r_escape, nullptr,
      nullptr, rf, vf, nact, nullptr);
}
extern "C" __global__ __launch_bounds__(64) void mms_axis3d_queue(
    const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ r0, const double* __restrict__ v0,
    double Bgrad, double mass, double dt, int n_steps, double r_escape, double* __restrict__ rf,
    double* __restrict__ vf, int* __restrict__ nact, MMQueue Q) {
  axis3d_queue_body<MMEvalSpec>(P, n, r0, v0, Bgrad, mass, dt, n_steps, r_escape, rf, vf, nact, Q, nullptr);
}
)";
    } else {
        emit_grad(o, P, L);
        o += R"(struct MMGradSpec {
  __device__ __forceinline__ static bool grad(double v, double z, double Bgrad, double& F, double (&J)[MM_ADJ_NJ]) {
    return mms_grad(v, z, Bgrad, F, J);
  }
};

extern "C" __global__ __launch_bounds__(64) void mms_tangent_queue(
    const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,
    double Bgrad, double mass, double dt, int n_steps, const int* __restrict__ n_up, double z_escape,
    double* __restrict__ zf, double* __restrict__ vf, int* __restrict__ n_done, double* __restrict__ sens,
    MMQueue Q) {
  tangent_queue_body<MMGradSpec, MMS_NG, MMS_NE>(P, n, z0, v0, Bgrad, mass, dt, n_steps, n_up, z_escape, zf, vf,
                                                  n_done, sens, Q);
}

extern "C" __global__ __launch_bounds__(64) void mms_tangent_grid(
    const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,
    double Bgrad, double mass, double dt, int n_steps, const int* __restrict__ n_up, double z_escape,
    double* __restrict__ zf, double* __restrict__ vf, int* __restrict__ n_done, double* __restrict__ sens) {
  const long i = (long)blockIdx.x * 64 + threadIdx.x;
  tangent_grid_body<MMGradSpec, MMS_NG, MMS_NE>(P, i, n, z0, v0, Bgrad, mass, dt, n_steps, n_up, z_escape, zf, vf,
                                                 n_done, sens);
}

extern "C" __global__ __launch_bounds__(64) void mms_vjp_segment(const MMRatePlanDev* __restrict__ P, VjpArgs A,
                                                                int seg) {
  vjp_segment_body<MMGradSpec, MMS_NG, MMS_NE>(P, A, seg, (long)blockIdx.x * 64 + threadIdx.x);
}

// F and the 9-entry Jacobian at scattered points (k_rate_grad's contract;
// NaN Jacobian where the reduced solve does not apply)
extern "C" __global__ __launch_bounds__(64) void mms_rate_grad_points(
    const MMRatePlanDev* __restrict__ P, const double* __restrict__ v, const double* __restrict__ z, long n,
    double Bgrad, double* __restrict__ F, double* __restrict__ Jout) {
  const long i = (long)blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  double f, J[MM_ADJ_NJ];
  if (!mms_grad(v[i], z[i], Bgrad, f, J)) {
    double fd, r;
    rate_eval_dense<MMS_NG, MMS_NE>(P, v[i], z[i], Bgrad, &fd, &r, nullptr);
    f = fd;
#pragma unroll
    for (int j = 0; j < MM_ADJ_NJ; j++) J[j] = __builtin_nan("");
  }
  if (F) F[i] = f;
#pragma unroll
  for (int j = 0; j < MM_ADJ_NJ; j++) Jout[i * MM_ADJ_NJ + j] = J[j];
}
)";
    }
    return o;
}

// Why a plan cannot have kernels of this kind, or null.
// The RF MOT's two phase plans (molmot/mot/simulator.py:49-90): unit beam
// directions, the same state counts, and for each plan a reduced solve with
// no dark ground state; packed or not (the RF phase plans drive a pair with
// one +z and one -z beam per frequency, outside the packed form).
static const char* spec_unsupported_rf(const MMRatePlanDev& P0, const MMRatePlanDev& P1) {
    if (P0.n_g != P1.n_g || P0.n_e != P1.n_e) return "the two phase plans differ in size";
    for (const MMRatePlanDev* P : {&P0, &P1}) {
        if (!P->unit_dirs) return "the beams of an RF phase plan must be along +-z";
        if (P->force_dense) return "an excited state has no decay channel (the reduced solve does not apply)";
        for (int g = 0; g < P->n_g; g++)
            if (coupled(*P, g).empty()) return "a ground state is dark (every point takes the dense solve)";
    }
    return nullptr;
}

// Source of the RF kernels: the first phase plan's evaluation at +Bgrad
// (mms_fast) and the second's at -Bgrad (mms_fast_b) per step, each with the
// dense fallback on its own plan's tables; 1D trajectories (with and without
// recoil kicks and the saved trajectory) and the Monte-Carlo temperature sums.
static std::string spec_source_rf(const MMRatePlanDev& P0, const MMRatePlanDev& P1) {
    std::string o = "// generated: plan-specialised RF MOT kernels (two phase plans)\n";
    o += "typedef __hip_internal::uint32_t uint32_t;\ntypedef __hip_internal::uint64_t uint64_t;\n";
    o += "#include \"mm_traj.h\"\n#include \"mm_vsq.h\"\n";
    emit(o, "#define MMS_NG %d\n#define MMS_NE %d\n", P0.n_g, P0.n_e);
    emit_fast(o, P0, spec_layout(P0), false, "mms_fast");
    emit_fast(o, P1, spec_layout(P1), false, "mms_fast_b");
    o += R"(template <int B>
struct MMEvalSpecRF {
  __device__ __forceinline__ static void eval(const MMRatePlanDev* __restrict__ P, double v, double z, double Bgrad,
                                              double& F, double& R, double*, int) {
    if (!(B ? mms_fast_b(v, z, Bgrad, F, R) : mms_fast(v, z, Bgrad, F, R))) {
      double Fd, Rd;
      rate_eval_dense<MMS_NG, MMS_NE>(P, v, z, Bgrad, &Fd, &Rd, nullptr);
      F = Fd;
      R = Rd;
    }
  }
};

#define MMS_RF_TRAJ(NAME, SAVE, KICKS)                                                                              \
  extern "C" __global__ __launch_bounds__(64) void NAME(                                                           \
      const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,    \
      double Bgrad, double mass, double dt, int n_steps, double z_escape, double* __restrict__ zt,                  \
      double* __restrict__ vt, double* __restrict__ zf, double* __restrict__ vf, int* __restrict__ nact,           \
      double kick_scale, uint32_t k0, uint32_t k1, uint32_t sid, const MMRatePlanDev* __restrict__ P2) {           \
    const long i = (long)blockIdx.x * 64 + threadIdx.x;                                                            \
    traj_grid_body<MMEvalSpecRF<0>, MMS_NG, MMS_NE, true, SAVE, KICKS, 64, true, MMEvalSpecRF<1>>(                 \
        P, i, n, z0, v0, Bgrad, mass, dt, n_steps, z_escape, zt, vt, zf, vf, nact, kick_scale, k0, k1, sid, P2,    \
        nullptr);                                                                                                   \
  }
MMS_RF_TRAJ(mms_rf_traj_k0, false, false)
MMS_RF_TRAJ(mms_rf_traj_k1, false, true)
MMS_RF_TRAJ(mms_rf_save_k0, true, false)
MMS_RF_TRAJ(mms_rf_save_k1, true, true)

extern "C" __global__ __launch_bounds__(64) void mms_rf_vsq(
    const MMRatePlanDev* __restrict__ P, long n, const double* __restrict__ z0, const double* __restrict__ v0,
    double Bgrad, double mass, double dt, int n_steps, int i_eq, double z_escape, double* __restrict__ vsq,
    long long* __restrict__ cnt, const MMRatePlanDev* __restrict__ P2) {
  vsq_grid_body<MMEvalSpecRF<0>, MMS_NG, MMS_NE, true, 64, true, MMEvalSpecRF<1>>(
      P, P2, (long)blockIdx.x * 64 + threadIdx.x, n, z0, v0, Bgrad, mass, dt, n_steps, i_eq, z_escape, vsq, cnt,
      nullptr);
}
)";
    return o;
}

static const char* spec_unsupported_kind(const MMRatePlanDev& P, int kind) {
    const char* why = spec_unsupported(P);
    if (why) return why;
    if (kind == MM_SPEC_GRADIENT && !spec_has_grad(P))
        return "the gradient kernels need one term of each of at most four frequency components per coupled pair";
    if (kind != MM_SPEC_FORWARD && kind != MM_SPEC_GRADIENT) return "kind must be 1 (forward) or 2 (gradient)";
    return nullptr;
}

long mm_spec_source_copy(const MMRatePlanDev& P, int kind, char* buf, long size) {
    const char* why = spec_unsupported_kind(P, kind);
    if (why) return mm_fail(MM_EUNSUPPORTED, "specialise: %s", why);
    const std::string src = spec_source(P, kind);
    if (buf && size > 0) {
        const long n = (long)src.size() < size - 1 ? (long)src.size() : size - 1;
        memcpy(buf, src.data(), (size_t)n);
        buf[n] = '\0';
    }
    return (long)src.size() + 1;
}

// ---------------------------------------------------------------------------
// compilation and module cache
// ---------------------------------------------------------------------------

static int rtc_compile(const std::string& src, const char* arch, std::vector<char>& code) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "mm_spec.hip", mm_rtc_n_headers, mm_rtc_header_srcs,
                            mm_rtc_header_names) != HIPRTC_SUCCESS)
        return mm_fail(MM_EHIP, "specialise: hiprtcCreateProgram failed");
    std::string off = std::string("--offload-arch=") + arch;
    // the library's own flags (molmot_amd/csrc/build.py): contraction off
    // (the detuning keeps the reference's rounding), no packed FP32.  And no
    // machine-level loop-invariant code motion: it hoists the scalar
    // materialisation of the ~180 table constants out of the step loop, where
    // they overflow the scalar registers and spill to vector-register lanes;
    // left in the loop they are re-materialised by scalar moves (forward
    // step: VALU instructions 1,564 -> 1,429, readlane / writelane 374 -> 230,
    // 100k x 30k launch 132.5 -> 123.0 ms, bit-identical trajectories)
    const char* opts[] = {off.c_str(), "-O3",     "-std=c++17", "-ffp-contract=off",
                          "-mllvm",    "-disable-machine-licm", "-Xclang",    "-target-feature",
                          "-Xclang",   "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
    if (r != HIPRTC_SUCCESS) {
        size_t n = 0;
        hiprtcGetProgramLogSize(prog, &n);
        std::string log(n + 1, '\0');
        if (n) hiprtcGetProgramLog(prog, &log[0]);
        hiprtcDestroyProgram(&prog);
        return mm_fail(MM_EHIP, "specialise: hipRTC compile failed (%s): %.400s", hiprtcGetErrorString(r),
                       log.c_str());
    }
    size_t n = 0;
    if (hiprtcGetCodeSize(prog, &n) != HIPRTC_SUCCESS || n == 0) {
        hiprtcDestroyProgram(&prog);
        return mm_fail(MM_EHIP, "specialise: hipRTC produced no code");
    }
    code.resize(n);
    const hiprtcResult rc = hiprtcGetCode(prog, code.data());
    hiprtcDestroyProgram(&prog);
    if (rc != HIPRTC_SUCCESS) return mm_fail(MM_EHIP, "specialise: hiprtcGetCode failed");
    return MM_OK;
}

// module for (device, source): from the cache or compiled and loaded now;
// called with mm_spec_mtx held and `device` current
static int spec_module(int device, int kind, const std::string& src, MMSpecModule** out) {
    auto key = std::make_pair(device, src);
    auto it = mm_spec_cache.find(key);
    if (it != mm_spec_cache.end()) {
        if (it->second->refs++ == 0)
            for (size_t i = 0; i < mm_spec_idle.size(); i++)
                if (mm_spec_idle[i] == it->second) {
                    mm_spec_idle.erase(mm_spec_idle.begin() + i);
                    break;
                }
        *out = it->second;
        return MM_OK;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return mm_hip_fail("specialise: hipGetDeviceProperties");
    std::vector<char> code;
    int rc = rtc_compile(src, prop.gcnArchName, code);
    if (rc != MM_OK) return rc;
    MMSpecModule* m = new (std::nothrow) MMSpecModule;
    if (!m) return mm_fail(MM_EINVAL, "specialise: host allocation failed");
    m->device = device;
    m->kind = kind;
    m->refs = 1;
    m->key = key;
    if (hipModuleLoadData(&m->mod, code.data()) != hipSuccess) {
        delete m;
        return mm_hip_fail("specialise: hipModuleLoadData");
    }
    static const char* fwd[15] = {"mms_traj_queue_k0", "mms_traj_queue_k1", "mms_traj_grid_k0", "mms_traj_grid_k1",
                                 "mms_vjp_forward",   "mms_rate_map",      "mms_rate_points",  "mms_vsq_grid",
                                 "mms_vsq_queue",     "mms_axis3d_grid",   "mms_axis3d_queue", "mms_traj_stats_k0",
                                 "mms_traj_stats_k1", "mms_traj_save_k0",  "mms_traj_save_k1"};
    static const char* grd[4] = {"mms_tangent_queue", "mms_tangent_grid", "mms_vjp_segment", "mms_rate_grad_points"};
    static const char* rf[5] = {"mms_rf_traj_k0", "mms_rf_traj_k1", "mms_rf_save_k0", "mms_rf_save_k1", "mms_rf_vsq"};
    const char* const* names = kind == MM_SPEC_FORWARD ? fwd : kind == MM_SPEC_GRADIENT ? grd : rf;
    const int nf = kind == MM_SPEC_FORWARD ? 15 : kind == MM_SPEC_GRADIENT ? 4 : 5;
    for (int i = 0; i < nf; i++)
        if (hipModuleGetFunction(&m->fn[i], m->mod, names[i]) != hipSuccess) {
            (void)hipModuleUnload(m->mod);
            delete m;
            return mm_hip_fail("specialise: hipModuleGetFunction");
        }
    mm_spec_cache[key] = m;
    *out = m;
    return MM_OK;
}

static void spec_unload(MMSpecModule* m) {
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur != m->device) (void)hipSetDevice(m->device);
    (void)hipModuleUnload(m->mod);
    if (cur >= 0 && cur != m->device) (void)hipSetDevice(cur);
    delete m;
}

// A module no plan references stays cached (plans built again from the same
// tables, e.g. one per call of a gradient routine, find it); beyond
// MM_SPEC_KEEP such modules the oldest is unloaded.
#define MM_SPEC_KEEP 8

static void spec_unref(MMSpecModule* m) {
    if (--m->refs > 0) return;
    mm_spec_idle.push_back(m);
    if (mm_spec_idle.size() > MM_SPEC_KEEP) {
        MMSpecModule* old = mm_spec_idle.front();
        mm_spec_idle.erase(mm_spec_idle.begin());
        mm_spec_cache.erase(old->key);
        spec_unload(old);
    }
}

void mm_spec_release(mm_rate_plan* pl) {
    if (!pl) return;
    std::lock_guard<std::mutex> lk(mm_spec_mtx);
    if (pl->spec) spec_unref(pl->spec);
    if (pl->spec_grad) spec_unref(pl->spec_grad);
    if (pl->spec_rf) spec_unref(pl->spec_rf);
    pl->spec = pl->spec_grad = pl->spec_rf = nullptr;
    pl->spec_rf_peer = nullptr;
}

static bool on_device(const MMSpecModule* m);

// The RF module of (p0, p1) is p0's while p1 is the plan it was built with
// (same handle and serial number: a plan built later at the same address
// carries another serial)
bool mm_spec_rf_ready(const mm_rate_plan* p0, const mm_rate_plan* p1) {
    return p0 && p1 && p0->spec_rf_peer == p1 && p0->spec_rf_serial == p1->serial && on_device(p0->spec_rf);
}

int mm_spec_launch_rf_traj(const mm_rate_plan* p0, const mm_rate_plan* p1, long n, const double* z0,
                           const double* v0, double B, double m, double dt, int ns, double ze, double* zt, double* vt,
                           double* zf, double* vf, int* na, double ks, uint32_t k0, uint32_t k1, uint32_t sid,
                           hipStream_t s) {
    const MMRatePlanDev* P = p0->dev;
    const MMRatePlanDev* P2 = p1->dev;
    void* args[] = {&P, &n, &z0, &v0, &B, &m, &dt, &ns, &ze, &zt, &vt, &zf, &vf, &na, &ks, &k0, &k1, &sid, &P2};
    const int f = (zt ? 2 : 0) + (ks > 0.0 ? 1 : 0);
    if (hipModuleLaunchKernel(p0->spec_rf->fn[f], (unsigned)((n + 63) / 64), 1, 1, 64, 1, 1, 0, s, args, nullptr) !=
        hipSuccess)
        return mm_hip_fail("launch mms_rf_traj");
    return MM_OK;
}

int mm_spec_launch_rf_vsq(const mm_rate_plan* p0, const mm_rate_plan* p1, long n, const double* z0,
                          const double* v0, double B, double m, double dt, int ns, int ieq, double ze, double* vsq,
                          long long* cnt, hipStream_t s) {
    const MMRatePlanDev* P = p0->dev;
    const MMRatePlanDev* P2 = p1->dev;
    void* args[] = {&P, &n, &z0, &v0, &B, &m, &dt, &ns, &ieq, &ze, &vsq, &cnt, &P2};
    if (hipModuleLaunchKernel(p0->spec_rf->fn[4], (unsigned)((n + 63) / 64), 1, 1, 64, 1, 1, 0, s, args, nullptr) !=
        hipSuccess)
        return mm_hip_fail("launch mms_rf_vsq");
    return MM_OK;
}

static bool on_device(const MMSpecModule* m) {
    int cur = -1;
    return m && hipGetDevice(&cur) == hipSuccess && cur == m->device;
}

bool mm_spec_ready(const mm_rate_plan* pl) { return pl && on_device(pl->spec); }
bool mm_spec_grad_ready(const mm_rate_plan* pl) { return pl && on_device(pl->spec_grad); }

int mm_spec_queue_slots(const mm_rate_plan* pl, bool kicks, MMQueueSlots& s) {
    return mm_queue_slots_fn(pl->spec->fn[kicks ? 1 : 0], s);
}

int mm_spec_tangent_slots(const mm_rate_plan* pl, MMQueueSlots& s) {
    return mm_queue_slots_fn(pl->spec_grad->fn[0], s);
}

int mm_spec_launch_grid(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                        double dt, int ns, double ze, double* zf, double* vf, int* na, double ks, uint32_t k0,
                        uint32_t k1, uint32_t sid, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    void* args[] = {&P, &n, &z0, &v0, &B, &m, &dt, &ns, &ze, &zf, &vf, &na, &ks, &k0, &k1, &sid};
    const unsigned grid = (unsigned)((n + 63) / 64);
    if (hipModuleLaunchKernel(pl->spec->fn[ks > 0.0 ? 3 : 2], grid, 1, 1, 64, 1, 1, 0, s, args, nullptr) !=
        hipSuccess)
        return mm_hip_fail("launch mms_traj_grid");
    return MM_OK;
}

int mm_spec_launch_grid_save(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                             double dt, int ns, double ze, double* zt, double* vt, double* zf, double* vf, int* na,
                             double ks, uint32_t k0, uint32_t k1, uint32_t sid, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    void* args[] = {&P, &n, &z0, &v0, &B, &m, &dt, &ns, &ze, &zt, &vt, &zf, &vf, &na, &ks, &k0, &k1, &sid};
    const unsigned grid = (unsigned)((n + 63) / 64);
    if (hipModuleLaunchKernel(pl->spec->fn[ks > 0.0 ? 14 : 13], grid, 1, 1, 64, 1, 1, 0, s, args, nullptr) !=
        hipSuccess)
        return mm_hip_fail("launch mms_traj_save");
    return MM_OK;
}

int mm_spec_launch_queue(const mm_rate_plan* pl, unsigned grid, size_t dyn, long n, const double* z0,
                         const double* v0, double B, double m, double dt, int ns, double ze, double* zf, double* vf,
                         int* na, double ks, uint32_t k0, uint32_t k1, uint32_t sid, const MMQueue& Q,
                         hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    MMQueue q = Q;
    void* args[] = {&P, &n, &z0, &v0, &B, &m, &dt, &ns, &ze, &zf, &vf, &na, &ks, &k0, &k1, &sid, &q};
    if (hipModuleLaunchKernel(pl->spec->fn[ks > 0.0 ? 1 : 0], grid, 1, 1, 64, 1, 1, (unsigned)dyn, s, args,
                              nullptr) != hipSuccess)
        return mm_hip_fail("launch mms_traj_queue");
    return MM_OK;
}

int mm_spec_launch_tangent_grid(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B,
                                double m, double dt, int ns, const int* n_up, double ze, double* zf, double* vf,
                                int* nd, double* sens, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    void* args[] = {&P, &n, &z0, &v0, &B, &m, &dt, &ns, &n_up, &ze, &zf, &vf, &nd, &sens};
    const unsigned grid = (unsigned)((n + 63) / 64);
    if (hipModuleLaunchKernel(pl->spec_grad->fn[1], grid, 1, 1, 64, 1, 1, 0, s, args, nullptr) != hipSuccess)
        return mm_hip_fail("launch mms_tangent_grid");
    return MM_OK;
}

int mm_spec_launch_tangent_queue(const mm_rate_plan* pl, unsigned grid, size_t dyn, long n, const double* z0,
                                 const double* v0, double B, double m, double dt, int ns, const int* n_up, double ze,
                                 double* zf, double* vf, int* nd, double* sens, const MMQueue& Q, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    MMQueue q = Q;
    void* args[] = {&P, &n, &z0, &v0, &B, &m, &dt, &ns, &n_up, &ze, &zf, &vf, &nd, &sens, &q};
    if (hipModuleLaunchKernel(pl->spec_grad->fn[0], grid, 1, 1, 64, 1, 1, (unsigned)dyn, s, args, nullptr) !=
        hipSuccess)
        return mm_hip_fail("launch mms_tangent_queue");
    return MM_OK;
}

int mm_spec_launch_vjp_forward(const mm_rate_plan* pl, const VjpArgs& A, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    VjpArgs a = A;
    void* args[] = {&P, &a};
    if (hipModuleLaunchKernel(pl->spec->fn[4], (unsigned)((A.n + 63) / 64), 1, 1, 64, 1, 1, 0, s, args, nullptr) !=
        hipSuccess)
        return mm_hip_fail("launch mms_vjp_forward");
    return MM_OK;
}

int mm_spec_launch_vjp_segment(const mm_rate_plan* pl, const VjpArgs& A, int seg, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    VjpArgs a = A;
    void* args[] = {&P, &a, &seg};
    if (hipModuleLaunchKernel(pl->spec_grad->fn[2], (unsigned)((A.n + 63) / 64), 1, 1, 64, 1, 1, 0, s, args,
                              nullptr) != hipSuccess)
        return mm_hip_fail("launch mms_vjp_segment");
    return MM_OK;
}

int mm_spec_launch_map(const mm_rate_plan* pl, const double* v, const double* z, int nz, long n, double B, double* F,
                       double* pop, double* R, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    void* args[] = {&P, &v, &z, &nz, &n, &B, &F, &pop, &R};
    if (hipModuleLaunchKernel(pl->spec->fn[5], (unsigned)((n + 63) / 64), 1, 1, 64, 1, 1, 0, s, args, nullptr) !=
        hipSuccess)
        return mm_hip_fail("launch mms_rate_map");
    return MM_OK;
}

int mm_spec_launch_points(const mm_rate_plan* pl, const double* v, const double* z, long n, double B, double* F,
                          double* pop, double* R, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    void* args[] = {&P, &v, &z, &n, &B, &F, &pop, &R};
    if (hipModuleLaunchKernel(pl->spec->fn[6], (unsigned)((n + 63) / 64), 1, 1, 64, 1, 1, 0, s, args, nullptr) !=
        hipSuccess)
        return mm_hip_fail("launch mms_rate_points");
    return MM_OK;
}

int mm_spec_launch_grad_points(const mm_rate_plan* pl, const double* v, const double* z, long n, double B, double* F,
                               double* J, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    void* args[] = {&P, &v, &z, &n, &B, &F, &J};
    if (hipModuleLaunchKernel(pl->spec_grad->fn[3], (unsigned)((n + 63) / 64), 1, 1, 64, 1, 1, 0, s, args,
                              nullptr) != hipSuccess)
        return mm_hip_fail("launch mms_rate_grad_points");
    return MM_OK;
}

int mm_spec_vsq_slots(const mm_rate_plan* pl, MMQueueSlots& s) { return mm_queue_slots_fn(pl->spec->fn[8], s); }

int mm_spec_launch_vsq_grid(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                            double dt, int ns, int ieq, double ze, double* vsq, long long* cnt, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    void* args[] = {&P, &n, &z0, &v0, &B, &m, &dt, &ns, &ieq, &ze, &vsq, &cnt};
    if (hipModuleLaunchKernel(pl->spec->fn[7], (unsigned)((n + 63) / 64), 1, 1, 64, 1, 1, 0, s, args, nullptr) !=
        hipSuccess)
        return mm_hip_fail("launch mms_vsq_grid");
    return MM_OK;
}

int mm_spec_launch_vsq_queue(const mm_rate_plan* pl, unsigned grid, size_t dyn, long n, const double* z0,
                             const double* v0, double B, double m, double dt, int ns, int ieq, double ze,
                             double* vsq, long long* cnt, double* zs, double* vs, const MMQueue& Q, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    MMQueue q = Q;
    void* args[] = {&P, &n, &z0, &v0, &B, &m, &dt, &ns, &ieq, &ze, &vsq, &cnt, &zs, &vs, &q};
    if (hipModuleLaunchKernel(pl->spec->fn[8], grid, 1, 1, 64, 1, 1, (unsigned)dyn, s, args, nullptr) != hipSuccess)
        return mm_hip_fail("launch mms_vsq_queue");
    return MM_OK;
}

int mm_spec_axis3d_slots(const mm_rate_plan* pl, MMQueueSlots& s) { return mm_queue_slots_fn(pl->spec->fn[10], s); }

int mm_spec_launch_axis3d_grid(const mm_rate_plan* pl, long n, const double* r0, const double* v0, double B, double m,
                               double dt, int ns, double re, double* rf, double* vf, int* na, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    void* args[] = {&P, &n, &r0, &v0, &B, &m, &dt, &ns, &re, &rf, &vf, &na};
    if (hipModuleLaunchKernel(pl->spec->fn[9], (unsigned)((4 * n + 63) / 64), 1, 1, 64, 1, 1, 0, s, args, nullptr) !=
        hipSuccess)
        return mm_hip_fail("launch mms_axis3d_grid");
    return MM_OK;
}

int mm_spec_launch_axis3d_queue(const mm_rate_plan* pl, unsigned grid, size_t dyn, long n, const double* r0,
                                const double* v0, double B, double m, double dt, int ns, double re, double* rf,
                                double* vf, int* na, const MMQueue& Q, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    MMQueue q = Q;
    void* args[] = {&P, &n, &r0, &v0, &B, &m, &dt, &ns, &re, &rf, &vf, &na, &q};
    if (hipModuleLaunchKernel(pl->spec->fn[10], grid, 1, 1, 64, 1, 1, (unsigned)dyn, s, args, nullptr) != hipSuccess)
        return mm_hip_fail("launch mms_axis3d_queue");
    return MM_OK;
}

int mm_spec_launch_stats(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                         double dt, int ns, double ze, double ra, const int* samp, int n_samp, double* part, double ks,
                         uint32_t k0, uint32_t k1, uint32_t sid, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    void* args[] = {&P, &n, &z0, &v0, &B, &m, &dt, &ns, &ze, &ra, &samp, &n_samp, &part, &ks, &k0, &k1, &sid};
    if (hipModuleLaunchKernel(pl->spec->fn[ks > 0.0 ? 12 : 11], (unsigned)((n + 63) / 64), 1, 1, 64, 1, 1, 0, s, args,
                              nullptr) != hipSuccess)
        return mm_hip_fail("launch mms_traj_stats");
    return MM_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int mm_rate_plan_specialize(mm_rate_plan* pl, int kinds) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "specialise: null plan");
    if (kinds < 0 || kinds > 3) return mm_fail(MM_EINVAL, "specialise: kinds is 0 (release) or a mask of 1, 2");
    if (!kinds) {
        mm_spec_release(pl);
        return MM_OK;
    }
    for (int kind = 1; kind <= 2; kind <<= 1) {
        if (!(kinds & kind)) continue;
        const char* why = spec_unsupported_kind(pl->host, kind);
        if (why) return mm_fail(MM_EUNSUPPORTED, "specialise: %s", why);
    }
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return mm_hip_fail("specialise: hipGetDevice");
    if (cur != pl->device && hipSetDevice(pl->device) != hipSuccess) return mm_hip_fail("specialise: hipSetDevice");
    int rc = MM_OK;
    for (int kind = 1; kind <= 2 && rc == MM_OK; kind <<= 1) {
        MMSpecModule** slot = kind == MM_SPEC_FORWARD ? &pl->spec : &pl->spec_grad;
        if (!(kinds & kind) || *slot) continue;
        const std::string src = spec_source(pl->host, kind);
        std::lock_guard<std::mutex> lk(mm_spec_mtx);
        rc = spec_module(pl->device, kind, src, slot);
    }
    if (cur != pl->device) (void)hipSetDevice(cur);
    return rc;
}

int mm_rate_plan_specialize_rf(mm_rate_plan* p0, const mm_rate_plan* p1) {
    mm_clear_error();
    if (!p0) return mm_fail(MM_EINVAL, "specialise_rf: null plan");
    {
        std::lock_guard<std::mutex> lk(mm_spec_mtx);
        if (p0->spec_rf) spec_unref(p0->spec_rf);
        p0->spec_rf = nullptr;
        p0->spec_rf_peer = nullptr;
    }
    if (!p1) return MM_OK;
    if (p0->device != p1->device) return mm_fail(MM_EINVAL, "specialise_rf: the phase plans are on different devices");
    const char* why = spec_unsupported_rf(p0->host, p1->host);
    if (why) return mm_fail(MM_EUNSUPPORTED, "specialise_rf: %s", why);
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return mm_hip_fail("specialise_rf: hipGetDevice");
    if (cur != p0->device && hipSetDevice(p0->device) != hipSuccess) return mm_hip_fail("specialise_rf: hipSetDevice");
    const std::string src = spec_source_rf(p0->host, p1->host);
    int rc;
    {
        std::lock_guard<std::mutex> lk(mm_spec_mtx);
        rc = spec_module(p0->device, MM_SPEC_RF, src, &p0->spec_rf);
        if (rc == MM_OK) {
            p0->spec_rf_peer = p1;
            p0->spec_rf_serial = p1->serial;
        }
    }
    if (cur != p0->device) (void)hipSetDevice(cur);
    return rc;
}

int mm_rate_plan_is_specialized_rf(const mm_rate_plan* p0, const mm_rate_plan* p1) {
    return mm_spec_rf_ready(p0, p1) ? 1 : 0;
}

long mm_rate_plan_spec_source_rf(const mm_rate_plan* p0, const mm_rate_plan* p1, char* buf, long size) {
    mm_clear_error();
    if (!p0 || !p1) return mm_fail(MM_EINVAL, "specialise_rf: null plan");
    const char* why = spec_unsupported_rf(p0->host, p1->host);
    if (why) return mm_fail(MM_EUNSUPPORTED, "specialise_rf: %s", why);
    const std::string src = spec_source_rf(p0->host, p1->host);
    if (buf && size > 0) {
        const long n = (long)src.size() < size - 1 ? (long)src.size() : size - 1;
        memcpy(buf, src.data(), (size_t)n);
        buf[n] = '\0';
    }
    return (long)src.size() + 1;
}

int mm_rate_plan_is_specialized(const mm_rate_plan* pl) {
    if (!pl) return 0;
    return (pl->spec ? MM_SPEC_FORWARD : 0) | (pl->spec_grad ? MM_SPEC_GRADIENT : 0);
}

long mm_rate_plan_spec_source(const mm_rate_plan* pl, int kind, char* buf, long size) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "specialise: null plan");
    return mm_spec_source_copy(pl->host, kind, buf, size);
}

long mm_rtc_compile(const char* src, const char* arch, char* code, long size) {
    mm_clear_error();
    if (!src || !arch) return mm_fail(MM_EINVAL, "rtc_compile: null source or architecture");
    std::vector<char> obj;
    const int rc = rtc_compile(std::string(src), arch, obj);
    if (rc != MM_OK) return rc;
    if (code && size > 0) memcpy(code, obj.data(), obj.size() < (size_t)size ? obj.size() : (size_t)size);
    return (long)obj.size();
}

}  // extern "C"
