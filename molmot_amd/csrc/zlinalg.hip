// Batched dense complex FP64 linear solves on gfx950, for the steady-state
// Lindblad problem (N^2 x N^2 Liouvillians, molmot_cuda/lindblad_cuda.py:
// 199-258, where the reference calls cuSOLVER) and the Floquet
// block-tridiagonal system (molmot/obe/floquet.py:200-300, where the
// reference calls a dense LAPACK solve).
//
// k_zgesv: one 256-thread workgroup per system, the matrix in HBM (row-major
// complex128), Gaussian elimination with partial pivoting (LAPACK's choice:
// the first row of largest |Re| + |Im| in the column, izamax), the right-hand
// sides carried along as extra columns of the elimination, then back
// substitution.  Per elimination step the workgroup finds the pivot
// (reduction through LDS), swaps the two rows across A and B, forms the
// multipliers into LDS, and updates the trailing columns: thread t owns
// columns t, t + 256, ... of [A | B], so every row of the update is one
// coalesced sweep.  Per system it reports the LAPACK info value (k + 1 for
// the first exactly-zero pivot) and min / max |U_kk| (|Re| + |Im|), from
// which the caller tells rank-deficient systems apart.
#include <hip/hip_runtime.h>

#include <math.h>

#include "../../include/molmot_hip.h"
#include "mm_errors.h"

#define ZL_BS 256
#define ZL_MAX_N 2048

struct zc {
    double r, i;
};

__device__ __forceinline__ zc zl_ld(const double* p) { return {p[0], p[1]}; }
__device__ __forceinline__ void zl_st(double* p, zc a) {
    p[0] = a.r;
    p[1] = a.i;
}
__device__ __forceinline__ double zl_abs1(zc a) { return fabs(a.r) + fabs(a.i); }
// x - l a
__device__ __forceinline__ zc zl_msub(zc x, zc l, zc a) {
    return {fma(-l.r, a.r, fma(l.i, a.i, x.r)), fma(-l.r, a.i, fma(-l.i, a.r, x.i))};
}
__device__ __forceinline__ zc zl_mul(zc a, zc b) {
    return {fma(a.r, b.r, -a.i * b.i), fma(a.r, b.i, a.i * b.r)};
}
// 1 / a by Smith's scaling (no overflow for any finite non-zero a)
__device__ __forceinline__ zc zl_inv(zc a) {
    if (fabs(a.r) >= fabs(a.i)) {
        const double q = a.i / a.r, d = a.r + a.i * q;
        return {1.0 / d, -q / d};
    }
    const double q = a.r / a.i, d = a.i + a.r * q;
    return {q / d, -1.0 / d};
}

__global__ __launch_bounds__(ZL_BS) void k_zgesv(int n, int nrhs, double* __restrict__ A, double* __restrict__ B,
                                                 int* __restrict__ info, double* __restrict__ umin,
                                                 double* __restrict__ umax) {
    __shared__ double lr[ZL_MAX_N], li[ZL_MAX_N];  // multipliers, then a column of U
    __shared__ double redv[ZL_BS / 64];
    __shared__ int redi[ZL_BS / 64];
    __shared__ int piv_s;
    const long sys = blockIdx.x;
    double* a = A + sys * (long)n * n * 2;
    double* b = B + sys * (long)n * nrhs * 2;
    const int t = threadIdx.x;
    const int ncol = n + nrhs;
    int inf = 0;
    double mn = INFINITY, mx = 0.0;
    for (int k = 0; k < n; k++) {
        // pivot: largest |Re| + |Im| in column k, first on ties
        double bv = -1.0;
        int bi = n;
        for (int i = k + t; i < n; i += ZL_BS) {
            const double v = zl_abs1(zl_ld(a + ((long)i * n + k) * 2));
            if (v > bv) {
                bv = v;
                bi = i;
            }
        }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) {
            const double ov = __shfl_xor(bv, o);
            const int oi = __shfl_xor(bi, o);
            if (ov > bv || (ov == bv && oi < bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if ((t & 63) == 0) {
            redv[t >> 6] = bv;
            redi[t >> 6] = bi;
        }
        __syncthreads();
        if (t == 0) {
            double v = redv[0];
            int p = redi[0];
            for (int w = 1; w < ZL_BS / 64; w++)
                if (redv[w] > v || (redv[w] == v && redi[w] < p)) {
                    v = redv[w];
                    p = redi[w];
                }
            piv_s = p;
        }
        __syncthreads();
        const int p = piv_s;
        // swap rows k and p of [A | B]
        if (p != k) {
            for (int c = t; c < ncol; c += ZL_BS) {
                double* x = (c < n) ? a + ((long)k * n + c) * 2 : b + ((long)k * nrhs + (c - n)) * 2;
                double* y = (c < n) ? a + ((long)p * n + c) * 2 : b + ((long)p * nrhs + (c - n)) * 2;
                const zc xv = zl_ld(x), yv = zl_ld(y);
                zl_st(x, yv);
                zl_st(y, xv);
            }
        }
        __syncthreads();
        const zc u = zl_ld(a + ((long)k * n + k) * 2);
        const double au = zl_abs1(u);
        mn = fmin(mn, au);
        mx = fmax(mx, au);
        const bool zero = (au == 0.0);
        if (zero && inf == 0) inf = k + 1;
        // multipliers l_i = a_ik / u into LDS and the L part of A
        const zc ui = zero ? zc{0.0, 0.0} : zl_inv(u);
        for (int i = k + 1 + t; i < n; i += ZL_BS) {
            double* x = a + ((long)i * n + k) * 2;
            const zc l = zero ? zc{0.0, 0.0} : zl_mul(zl_ld(x), ui);
            zl_st(x, l);
            lr[i] = l.r;
            li[i] = l.i;
        }
        __syncthreads();
        // trailing update of columns k+1 .. n-1 of A and every column of B
        for (int c = k + 1 + t; c < ncol; c += ZL_BS) {
            double* col;
            long stride;
            if (c < n) {
                col = a + (long)c * 2;
                stride = (long)n * 2;
            } else {
                col = b + (long)(c - n) * 2;
                stride = (long)nrhs * 2;
            }
            const zc pk = zl_ld(col + k * stride);
            if (pk.r == 0.0 && pk.i == 0.0) continue;
            for (int i = k + 1; i < n; i++) {
                double* x = col + i * stride;
                zl_st(x, zl_msub(zl_ld(x), zc{lr[i], li[i]}, pk));
            }
        }
        __syncthreads();
    }
    // back substitution U X = B', column k of U staged in LDS
    for (int k = n - 1; k >= 0; k--) {
        for (int i = t; i < k; i += ZL_BS) {
            const zc x = zl_ld(a + ((long)i * n + k) * 2);
            lr[i] = x.r;
            li[i] = x.i;
        }
        const zc u = zl_ld(a + ((long)k * n + k) * 2);
        const zc ui = (u.r == 0.0 && u.i == 0.0) ? zc{0.0, 0.0} : zl_inv(u);
        __syncthreads();
        for (int r = t; r < nrhs; r += ZL_BS) {
            double* col = b + (long)r * 2;
            const long stride = (long)nrhs * 2;
            const zc xk = zl_mul(zl_ld(col + k * stride), ui);
            zl_st(col + k * stride, xk);
            for (int i = 0; i < k; i++) {
                double* x = col + i * stride;
                zl_st(x, zl_msub(zl_ld(x), zc{lr[i], li[i]}, xk));
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        if (info) info[sys] = inf;
        if (umin) umin[sys] = mn;
        if (umax) umax[sys] = mx;
    }
}

extern "C" int mm_zgesv_batched(int n, int nrhs, long batch, double* A, double* B, int* info, double* umin,
                                double* umax, void* stream) {
    mm_clear_error();
    if (n < 1 || n > ZL_MAX_N || nrhs < 1 || batch < 0)
        return mm_fail(MM_EINVAL, "zgesv_batched: 1 <= n <= %d, nrhs >= 1 and batch >= 0 required", ZL_MAX_N);
    if (batch == 0) return MM_OK;
    if (!A || !B) return mm_fail(MM_EINVAL, "zgesv_batched: null pointer");
    if (batch > 0x7fffffffl) return mm_fail(MM_EINVAL, "zgesv_batched: batch too large");
    hipLaunchKernelGGL(k_zgesv, dim3((unsigned)batch), dim3(ZL_BS), 0, (hipStream_t)stream, n, nrhs, A, B, info,
                       umin, umax);
    return mm_check_launch("k_zgesv");
}
