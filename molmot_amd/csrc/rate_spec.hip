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
// is cached per (device, generated source) for the life of the process and
// shared by every plan with the same tables.
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

// the device headers, embedded by molmot_amd/csrc/build.py
extern const char* const mm_rtc_header_names[];
extern const char* const mm_rtc_header_srcs[];
extern const int mm_rtc_n_headers;

struct MMSpecModule {
    int device;
    hipModule_t mod;
    hipFunction_t queue[2], grid[2];  // [kicks]
    int refs;
    std::pair<int, std::string> key;
};

static std::mutex mm_spec_mtx;
static std::map<std::pair<int, std::string>, MMSpecModule*> mm_spec_cache;

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

// Source of the specialised evaluator and its four kernels.
static std::string spec_source(const MMRatePlanDev& P) {
    const int NG = P.n_g, NE = P.n_e;
    std::string o;
    o += "// generated: plan-specialised 1D trajectory kernels\n";
    // hipRTC declares the fixed-width integer types in its own namespace
    o += "typedef __hip_internal::uint32_t uint32_t;\ntypedef __hip_internal::uint64_t uint64_t;\n";
    o += "#include \"mm_traj.h\"\n";
    emit(o, "#define MMS_NG %d\n#define MMS_NE %d\n", NG, NE);

    // distinct beam frequencies of the +z slots (0, 1) and -z slots (2, 3)
    std::vector<double> fp, fm;
    std::vector<int> slot(NG * 4 * 4, -1);
    for (int g = 0; g < NG; g++)
        for (int e = 0; e < NE; e++) {
            if (!(P.pair_mask[g] & (1u << e))) continue;
            const MMPairRec& Q = P.prec[g * 4 + e];
            for (int j = 0; j < 4; j++)
                slot[(g * 4 + e) * 4 + j] = j < 2 ? distinct_index(fp, Q.f[j]) : distinct_index(fm, Q.f[j]);
        }

    o += "__device__ __forceinline__ bool mms_fast(double v, double z, double Bgrad, double& F, double& R) {\n";
    o += "  const double Bg = Bgrad * z * 1e4;\n";
    emit(o, "  const double dop = mm_div_rn(%s * v, MM_TWO_PI, MM_INV_TWO_PI);\n", lit(P.mk).c_str());
    for (int e = 0; e < NE; e++) {
        const double E = P.E[NG + e], zd = P.zd[NG + e];
        if (zd == 0.0)
            emit(o, "  const double Ee%d = %s;\n", e, lit(E).c_str());
        else
            emit(o, "  const double Ee%d = %s + mm_div_rn(Bg * %s * %s, MM_TWO_PI, MM_INV_TWO_PI);\n", e,
                 lit(E).c_str(), lit(zd).c_str(), lit(P.Gamma).c_str());
    }
    for (size_t i = 0; i < fp.size(); i++) emit(o, "  const double fp%zu = %s + dop;\n", i, lit(fp[i]).c_str());
    for (size_t i = 0; i < fm.size(); i++) emit(o, "  const double fm%zu = %s - dop;\n", i, lit(fm[i]).c_str());
    o += "  double K[MMS_NE][MMS_NE];\n";
    o += "#pragma unroll\n  for (int e = 0; e < MMS_NE; e++)\n#pragma unroll\n"
         "    for (int f = 0; f < MMS_NE; f++) K[e][f] = 0.0;\n";
    o += "  bool ok = true;\n";
    for (int g = 0; g < NG; g++) emit(o, "  double D%d, G%d;\n", g, g);
    const std::string gq = lit(P.gq);
    for (int g = 0; g < NG; g++) {
        emit(o, "  { // ground state %d\n", g);
        const double E = P.E[g], zd = P.zd[g];
        if (zd == 0.0)
            emit(o, "    const double Eg = %s;\n", lit(E).c_str());
        else
            emit(o, "    const double Eg = %s + mm_div_rn(Bg * %s * %s, MM_TWO_PI, MM_INV_TWO_PI);\n",
                 lit(E).c_str(), lit(zd).c_str(), lit(P.Gamma).c_str());
        std::vector<int> es;
        for (int e = 0; e < NE; e++)
            if (P.pair_mask[g] & (1u << e)) es.push_back(e);
        for (int e : es) {
            const MMPairRec& Q = P.prec[g * 4 + e];
            const int* sl = &slot[(g * 4 + e) * 4];
            emit(o, "    double ps%d, pt%d;\n    {\n      const double om = Ee%d - Eg;\n", e, e, e);
            emit(o, "      const double da = fp%d - om, db = fp%d - om, dc = fm%d - om, dd = fm%d - om;\n", sl[0],
                 sl[1], sl[2], sl[3]);
            emit(o, "      const double xa = fma(da, da, %s), xb = fma(db, db, %s);\n", gq.c_str(), gq.c_str());
            emit(o, "      const double xc = fma(dc, dc, %s), xd = fma(dd, dd, %s);\n", gq.c_str(), gq.c_str());
            o += "      const double pab = xa * xb, pcd = xc * xd;\n";
            o += "      const double r = mm_rcp1(pab * pcd);\n";
            emit(o, "      const double np = fma(%s, xb, %s * xa) * pcd;\n", lit(Q.c2[0]).c_str(),
                 lit(Q.c2[1]).c_str());
            emit(o, "      const double nm = fma(%s, xd, %s * xc) * pab;\n", lit(Q.c2[2]).c_str(),
                 lit(Q.c2[3]).c_str());
            emit(o, "      ps%d = (np + nm) * r;\n      pt%d = (np - nm) * r;\n    }\n", e, e);
        }
        // Sg, Tg: the generic kernel's sums over all four pairs, whose empty
        // pairs add exact zeros
        std::string sg = "ps" + std::to_string(es[0]), tg = "pt" + std::to_string(es[0]);
        for (size_t k = 1; k < es.size(); k++) {
            sg = "(" + sg + " + ps" + std::to_string(es[k]) + ")";
            tg = "(" + tg + " + pt" + std::to_string(es[k]) + ")";
        }
        emit(o, "    const double Sg = %s, Tg = %s;\n", sg.c_str(), tg.c_str());
        o += "    ok = ok && (Sg > 0.0);\n";
        o += "    const double s1 = 1.0 + Sg;\n    const double rr = mm_rcp(Sg * s1);\n";
        o += "    const double rS = rr * s1, sat = rr * Sg;\n";
        for (int e : es) {
            bool any = false;
            for (int f = 0; f < NE; f++) any = any || (f != e && P.BR[f][g] != 0.0);
            if (!any) continue;
            emit(o, "    {\n      const double a = ps%d * rS;\n", e);
            for (int f = 0; f < NE; f++)
                if (f != e && P.BR[f][g] != 0.0)
                    emit(o, "      K[%d][%d] = fma(a, %s, K[%d][%d]);\n", e, f, lit(P.BR[f][g]).c_str(), e, f);
            o += "    }\n";
        }
        emit(o, "    D%d = (rS * s1) * %s;\n", g, lit(P.inv_gh_geff).c_str());
        emit(o, "    G%d = %s * sat * Tg;\n  }\n", g, lit(P.fk_gh_geff).c_str());
    }
    o += "  double y[MMS_NE];\n  mm_gth_stationary<MMS_NE>(K, ok, y);\n";
    o += "  double Z = 0.0, Fs = 0.0, Rs = 0.0;\n";
    for (int g = 0; g < NG; g++) {
        o += "  {\n    double u = 0.0;\n";
        for (int f = 0; f < NE; f++)
            if (P.BR[f][g] != 0.0) emit(o, "    u = fma(y[%d], %s, u);\n", f, lit(P.BR[f][g]).c_str());
        emit(o, "    const double q = u * D%d;\n    Z += q;\n    Fs = fma(q, G%d, Fs);\n    Rs += u;\n  }\n", g, g);
    }
    o += "  const double rZ = mm_rcp(Z);\n  F = Fs * rZ;\n  R = Rs * rZ;\n  return ok;\n}\n\n";

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
MMS_QUEUE(mms_traj_queue_k0, false)
MMS_QUEUE(mms_traj_queue_k1, true)
MMS_GRID(mms_traj_grid_k0, false)
MMS_GRID(mms_traj_grid_k1, true)
)";
    return o;
}

long mm_spec_source_copy(const MMRatePlanDev& P, char* buf, long size) {
    const char* why = spec_unsupported(P);
    if (why) return mm_fail(MM_EUNSUPPORTED, "specialise: %s", why);
    const std::string src = spec_source(P);
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
    // (the detuning keeps the reference's rounding), no packed FP32
    const char* opts[] = {off.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops"};
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
static int spec_module(int device, const std::string& src, MMSpecModule** out) {
    auto key = std::make_pair(device, src);
    auto it = mm_spec_cache.find(key);
    if (it != mm_spec_cache.end()) {
        it->second->refs++;
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
    m->refs = 1;
    m->key = key;
    if (hipModuleLoadData(&m->mod, code.data()) != hipSuccess) {
        delete m;
        return mm_hip_fail("specialise: hipModuleLoadData");
    }
    const char* names[4] = {"mms_traj_queue_k0", "mms_traj_queue_k1", "mms_traj_grid_k0", "mms_traj_grid_k1"};
    hipFunction_t* fns[4] = {&m->queue[0], &m->queue[1], &m->grid[0], &m->grid[1]};
    for (int i = 0; i < 4; i++)
        if (hipModuleGetFunction(fns[i], m->mod, names[i]) != hipSuccess) {
            hipModuleUnload(m->mod);
            delete m;
            return mm_hip_fail("specialise: hipModuleGetFunction");
        }
    mm_spec_cache[key] = m;
    *out = m;
    return MM_OK;
}

static void spec_unref(MMSpecModule* m) {
    if (--m->refs > 0) return;
    mm_spec_cache.erase(m->key);
    int cur = -1;
    hipGetDevice(&cur);
    if (cur != m->device) hipSetDevice(m->device);
    hipModuleUnload(m->mod);
    if (cur >= 0 && cur != m->device) hipSetDevice(cur);
    delete m;
}

void mm_spec_release(mm_rate_plan* pl) {
    if (!pl || !pl->spec) return;
    std::lock_guard<std::mutex> lk(mm_spec_mtx);
    spec_unref(pl->spec);
    pl->spec = nullptr;
}

bool mm_spec_ready(const mm_rate_plan* pl) {
    if (!pl || !pl->spec) return false;
    int cur = -1;
    return hipGetDevice(&cur) == hipSuccess && cur == pl->spec->device;
}

int mm_spec_queue_slots(const mm_rate_plan* pl, bool kicks, MMQueueSlots& s) {
    return mm_queue_slots_fn(pl->spec->queue[kicks ? 1 : 0], s);
}

int mm_spec_launch_grid(const mm_rate_plan* pl, long n, const double* z0, const double* v0, double B, double m,
                        double dt, int ns, double ze, double* zf, double* vf, int* na, double ks, uint32_t k0,
                        uint32_t k1, uint32_t sid, hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    void* args[] = {&P, &n, &z0, &v0, &B, &m, &dt, &ns, &ze, &zf, &vf, &na, &ks, &k0, &k1, &sid};
    const unsigned grid = (unsigned)((n + 63) / 64);
    if (hipModuleLaunchKernel(pl->spec->grid[ks > 0.0 ? 1 : 0], grid, 1, 1, 64, 1, 1, 0, s, args, nullptr) !=
        hipSuccess)
        return mm_hip_fail("launch mms_traj_grid");
    return MM_OK;
}

int mm_spec_launch_queue(const mm_rate_plan* pl, unsigned grid, size_t dyn, long n, const double* z0,
                         const double* v0, double B, double m, double dt, int ns, double ze, double* zf, double* vf,
                         int* na, double ks, uint32_t k0, uint32_t k1, uint32_t sid, const MMQueue& Q,
                         hipStream_t s) {
    const MMRatePlanDev* P = pl->dev;
    MMQueue q = Q;
    void* args[] = {&P, &n, &z0, &v0, &B, &m, &dt, &ns, &ze, &zf, &vf, &na, &ks, &k0, &k1, &sid, &q};
    if (hipModuleLaunchKernel(pl->spec->queue[ks > 0.0 ? 1 : 0], grid, 1, 1, 64, 1, 1, (unsigned)dyn, s, args,
                              nullptr) != hipSuccess)
        return mm_hip_fail("launch mms_traj_queue");
    return MM_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

int mm_rate_plan_specialize(mm_rate_plan* pl, int enable) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "specialise: null plan");
    if (!enable) {
        mm_spec_release(pl);
        return MM_OK;
    }
    if (pl->spec) return MM_OK;
    const char* why = spec_unsupported(pl->host);
    if (why) return mm_fail(MM_EUNSUPPORTED, "specialise: %s", why);
    const std::string src = spec_source(pl->host);
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return mm_hip_fail("specialise: hipGetDevice");
    if (cur != pl->device && hipSetDevice(pl->device) != hipSuccess) return mm_hip_fail("specialise: hipSetDevice");
    int rc;
    {
        std::lock_guard<std::mutex> lk(mm_spec_mtx);
        rc = spec_module(pl->device, src, &pl->spec);
    }
    if (cur != pl->device) hipSetDevice(cur);
    return rc;
}

int mm_rate_plan_is_specialized(const mm_rate_plan* pl) { return pl && pl->spec ? 1 : 0; }

long mm_rate_plan_spec_source(const mm_rate_plan* pl, char* buf, long size) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "specialise: null plan");
    return mm_spec_source_copy(pl->host, buf, size);
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
