// gfx950 kernel and C ABI for the per-point force gradient (rate_adjoint.h).
#include <hip/hip_runtime.h>

#include <math.h>

#include "../../include/molmot_hip.h"
#include "mm_errors.h"
#include "mm_plan_host.h"
#include "rate_adjoint.h"

template <int NG, int NE, bool UNIT, int BS>
__global__ __launch_bounds__(BS) void k_rate_grad(const MMRatePlanDev* __restrict__ P, const double* __restrict__ v,
                                                  const double* __restrict__ z, long n, double Bgrad,
                                                  double* __restrict__ F, double* __restrict__ Jout) {
    __shared__ double lds[mm_adj_slab<NG, NE>() * BS];
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double f, J[MM_ADJ_NJ];
    const bool ok = rate_grad_fast<NG, NE, UNIT>(P, v[i], z[i], Bgrad, f, J, lds + threadIdx.x, BS);
    if (!ok) {
        // a dark ground state or a reducible excited chain: the force is
        // still defined (dense path) but this gradient is not provided
        double r;
        rate_eval_dense<NG, NE>(P, v[i], z[i], Bgrad, &f, &r, nullptr);
#pragma unroll
        for (int j = 0; j < MM_ADJ_NJ; j++) J[j] = __builtin_nan("");
    }
    if (F) F[i] = f;
#pragma unroll
    for (int j = 0; j < MM_ADJ_NJ; j++) Jout[i * MM_ADJ_NJ + j] = J[j];
}

template <int NG>
constexpr int grad_block() {
    return 64;
}

template <int NG, int NE>
static int launch_grad(const mm_rate_plan* pl, long n, const double* v, const double* z, double B, double* F,
                       double* J, hipStream_t s) {
    constexpr int block = grad_block<NG>();
    if (pl->host.unit_dirs)
        hipLaunchKernelGGL((k_rate_grad<NG, NE, true, block>), dim3(grid_for(n, block)), dim3(block), 0, s, pl->dev,
                           v, z, n, B, F, J);
    else
        hipLaunchKernelGGL((k_rate_grad<NG, NE, false, block>), dim3(grid_for(n, block)), dim3(block), 0, s,
                           pl->dev, v, z, n, B, F, J);
    return mm_check_launch("k_rate_grad");
}

// Sizes with a gradient kernel (the LDS slab of 36+4 would not fit).
#define MM_GRAD_SIZE_LIST(X) \
    X(12, 4)                 \
    X(2, 1)                  \
    X(3, 1)                  \
    X(4, 2)

extern "C" int mm_rate_eq_grad_points(const mm_rate_plan* pl, long n, const double* v, const double* z, double B,
                                      double* F, double* J, void* stream) {
    mm_clear_error();
    if (!pl) return mm_fail(MM_EINVAL, "rate_eq_grad_points: null plan");
    if (n < 0) return mm_fail(MM_EINVAL, "rate_eq_grad_points: negative size");
    if (n == 0) return MM_OK;
    if (!v || !z || !J) return mm_fail(MM_EINVAL, "rate_eq_grad_points: null pointer");
    if (pl->host.n_comp > MM_ADJ_NC)
        return mm_fail(MM_EUNSUPPORTED, "rate_eq_grad_points: at most %d frequency components (plan has %d)",
                       MM_ADJ_NC, pl->host.n_comp);
    if (pl->host.force_dense)
        return mm_fail(MM_EUNSUPPORTED, "rate_eq_grad_points: an excited state has no decay channel");
    const int ng = pl->host.n_g, ne = pl->host.n_e;
#define MM_GR(G, E) \
    if (ng == G && ne == E) return launch_grad<G, E>(pl, n, v, z, B, F, J, (hipStream_t)stream);
    MM_GRAD_SIZE_LIST(MM_GR)
#undef MM_GR
    return mm_fail(MM_EUNSUPPORTED, "rate_eq_grad_points: no gradient kernel for (%d, %d)", ng, ne);
}
