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
// multipliers into LDS, and updates the trailing block of [A | B] (see the
// item scheme below).  Per system it reports the LAPACK info value (k + 1 for
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

// The elimination's inner updates run over (row, column) items flattened
// row-major and dealt round-robin to the workgroup's threads: consecutive
// threads touch consecutive columns of a row (coalesced), each thread owns
// many independent items (loads of several items in flight: ZL_U per
// iteration, all loads before any store), and one right-hand side (the
// steady state) still spreads its rows over the whole workgroup.  LDS
// (dynamic): the multipliers and U column (n), the pivot row of [A | B]
// (n + nrhs), and the current row of X (nrhs), complex.
#define ZL_U 4

struct ZlIter {
    int i, c;  // current item (row offset, column offset)
    int di, dc, nc;
    __device__ __forceinline__ void init(int t, int ncols) {
        nc = ncols;
        i = t / nc;
        c = t - i * nc;
        di = ZL_BS / nc;
        dc = ZL_BS - di * nc;
    }
    __device__ __forceinline__ void next() {
        i += di;
        c += dc;
        if (c >= nc) {
            c -= nc;
            i += 1;
        }
    }
};

__global__ __launch_bounds__(ZL_BS) void k_zgesv(int n, int nrhs, double* __restrict__ A, double* __restrict__ B,
                                                 int* __restrict__ info, double* __restrict__ umin,
                                                 double* __restrict__ umax) {
    extern __shared__ double zl_lds[];
    double* const mr = zl_lds;               // multipliers / U column, n complex
    double* const pr = zl_lds + 2 * n;       // pivot row of [A | B], n + nrhs complex
    double* const xr = pr + 2 * (n + nrhs);  // row of X, nrhs complex
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
    auto elem = [&](int row, int col) -> double* {
        return (col < n) ? a + ((long)row * n + col) * 2 : b + ((long)row * nrhs + (col - n)) * 2;
    };
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
        // swap rows k and p of [A | B]; the new pivot row (columns > k) to LDS
        for (int c = t; c < ncol; c += ZL_BS) {
            double* x = elem(k, c);
            zc xv = zl_ld(x);
            if (p != k) {
                double* y = elem(p, c);
                const zc yv = zl_ld(y);
                zl_st(x, yv);
                zl_st(y, xv);
                xv = yv;
            }
            pr[2 * c] = xv.r;
            pr[2 * c + 1] = xv.i;
        }
        __syncthreads();
        const zc u = {pr[2 * k], pr[2 * k + 1]};
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
            mr[2 * i] = l.r;
            mr[2 * i + 1] = l.i;
        }
        __syncthreads();
        // trailing update: rows k+1 .. n-1, columns k+1 .. ncol-1 of [A | B]
        const int nr = n - k - 1, nc = ncol - k - 1;
        if (nr > 0) {
            ZlIter it;
            it.init(t, nc);
            while (it.i < nr) {
                int ri[ZL_U], ci[ZL_U];
                zc xv[ZL_U];
#pragma unroll
                for (int u2 = 0; u2 < ZL_U; u2++) {
                    ri[u2] = it.i;
                    ci[u2] = it.c;
                    if (it.i < nr) xv[u2] = zl_ld(elem(k + 1 + it.i, k + 1 + it.c));
                    it.next();
                }
#pragma unroll
                for (int u2 = 0; u2 < ZL_U; u2++) {
                    if (ri[u2] < nr) {
                        const int i = k + 1 + ri[u2], c = k + 1 + ci[u2];
                        const zc l = {mr[2 * i], mr[2 * i + 1]};
                        const zc pk = {pr[2 * c], pr[2 * c + 1]};
                        zl_st(elem(i, c), zl_msub(xv[u2], l, pk));
                    }
                }
            }
        }
        __syncthreads();
    }
    // back substitution U X = B': row k of X to LDS, column k of U to LDS,
    // then every (row < k, rhs) item updated
    for (int k = n - 1; k >= 0; k--) {
        const zc u = zl_ld(a + ((long)k * n + k) * 2);
        const zc ui = (u.r == 0.0 && u.i == 0.0) ? zc{0.0, 0.0} : zl_inv(u);
        for (int r = t; r < nrhs; r += ZL_BS) {
            double* x = b + ((long)k * nrhs + r) * 2;
            const zc xk = zl_mul(zl_ld(x), ui);
            zl_st(x, xk);
            xr[2 * r] = xk.r;
            xr[2 * r + 1] = xk.i;
        }
        for (int i = t; i < k; i += ZL_BS) {
            const zc x = zl_ld(a + ((long)i * n + k) * 2);
            mr[2 * i] = x.r;
            mr[2 * i + 1] = x.i;
        }
        __syncthreads();
        if (k > 0) {
            ZlIter it;
            it.init(t, nrhs);
            while (it.i < k) {
                int ri[ZL_U], ci[ZL_U];
                zc xv[ZL_U];
#pragma unroll
                for (int u2 = 0; u2 < ZL_U; u2++) {
                    ri[u2] = it.i;
                    ci[u2] = it.c;
                    if (it.i < k) xv[u2] = zl_ld(b + ((long)it.i * nrhs + it.c) * 2);
                    it.next();
                }
#pragma unroll
                for (int u2 = 0; u2 < ZL_U; u2++) {
                    if (ri[u2] < k) {
                        const int i = ri[u2], r = ci[u2];
                        const zc l = {mr[2 * i], mr[2 * i + 1]};
                        const zc xk = {xr[2 * r], xr[2 * r + 1]};
                        zl_st(b + ((long)i * nrhs + r) * 2, zl_msub(xv[u2], l, xk));
                    }
                }
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

// ---------------------------------------------------------------------------
// Blocked form for n <= 256 (the N^2 x N^2 Liouvillians and Floquet blocks
// of up to 16 levels), panels of ZB = 16 columns.  Per panel: the panel
// (rows kb .. n-1) is factorised with partial pivoting in LDS, written back,
// its row swaps applied to the columns right of it and to B, the 16 rows of
// U12 = L11^-1 A12 formed, and the trailing block of [A | B] updated once,
// A22 -= L21 U12, with L21 read from LDS (broadcast) and a wavefront's 64
// consecutive columns of U12 held in registers: 16 complex multiply-adds per
// element loaded and stored, where the unblocked form does one.  Back
// substitution likewise by 16-row blocks.  LDS (dynamic): the panel (n x 16
// complex) and 16 rows of X (16 x nrhs complex).
// ---------------------------------------------------------------------------
#define ZB 16
// rows per thread in flight in the trailing update
#ifndef ZL_ILP
#define ZL_ILP 8
#endif
#define ZB_MAX_N 256

__global__ __launch_bounds__(ZL_BS) void k_zgesv_blk(int n, int nrhs, double* __restrict__ A,
                                                     double* __restrict__ B, int* __restrict__ info,
                                                     double* __restrict__ umin, double* __restrict__ umax) {
    extern __shared__ double zl_lds[];
    double* const P = zl_lds;                 // panel: [row][ZB] complex, n rows
    double* const XB = zl_lds + 2 * n * ZB;   // block of X: [ZB][nrhs] complex
    __shared__ double redv[ZL_BS / 64];
    __shared__ int redi[ZL_BS / 64];
    __shared__ int piv[ZB];
    const long sys = blockIdx.x;
    double* a = A + sys * (long)n * n * 2;
    double* b = B + sys * (long)n * nrhs * 2;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int ncol = n + nrhs;
    int inf = 0;
    double mn = INFINITY, mx = 0.0;
    auto elem = [&](int row, int col) -> double* {
        return (col < n) ? a + ((long)row * n + col) * 2 : b + ((long)row * nrhs + (col - n)) * 2;
    };
    auto pel = [&](int r, int j) -> double* { return P + (r * ZB + j) * 2; };
    for (int kb = 0; kb < n; kb += ZB) {
        const int w = min(ZB, n - kb), m = n - kb;
        // (a) panel to LDS
        for (int q = t; q < m * w; q += ZL_BS) {
            const int r = q / w, j = q - (q / w) * w;
            const zc x = zl_ld(a + ((long)(kb + r) * n + kb + j) * 2);
            zl_st(pel(r, j), x);
        }
        __syncthreads();
        // (b) factorise the panel in LDS
        for (int j = 0; j < w; j++) {
            double bv = -1.0;
            int bi = m;
            for (int r = j + t; r < m; r += ZL_BS) {
                const double v = zl_abs1(zl_ld(pel(r, j)));
                if (v > bv) {
                    bv = v;
                    bi = r;
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
            if (lane == 0) {
                redv[wv] = bv;
                redi[wv] = bi;
            }
            __syncthreads();
            if (t == 0) {
                double v = redv[0];
                int p = redi[0];
                for (int q = 1; q < ZL_BS / 64; q++)
                    if (redv[q] > v || (redv[q] == v && redi[q] < p)) {
                        v = redv[q];
                        p = redi[q];
                    }
                piv[j] = p;
            }
            __syncthreads();
            const int p = piv[j];
            if (p != j && t < w) {
                const zc x = zl_ld(pel(j, t)), y = zl_ld(pel(p, t));
                zl_st(pel(j, t), y);
                zl_st(pel(p, t), x);
            }
            __syncthreads();
            const zc u = zl_ld(pel(j, j));
            const double au = zl_abs1(u);
            mn = fmin(mn, au);
            mx = fmax(mx, au);
            const bool zero = (au == 0.0);
            if (zero && inf == 0) inf = kb + j + 1;
            const zc ui = zero ? zc{0.0, 0.0} : zl_inv(u);
            for (int r = j + 1 + t; r < m; r += ZL_BS) zl_st(pel(r, j), zero ? zc{0.0, 0.0} : zl_mul(zl_ld(pel(r, j)), ui));
            __syncthreads();
            const int wr = w - j - 1;
            if (wr > 0)
                for (int q = t; q < (m - j - 1) * wr; q += ZL_BS) {
                    const int r = j + 1 + q / wr, c = j + 1 + (q - (q / wr) * wr);
                    zl_st(pel(r, c), zl_msub(zl_ld(pel(r, c)), zl_ld(pel(r, j)), zl_ld(pel(j, c))));
                }
            __syncthreads();
        }
        // (c) panel back to A
        for (int q = t; q < m * w; q += ZL_BS) {
            const int r = q / w, j = q - (q / w) * w;
            zl_st(a + ((long)(kb + r) * n + kb + j) * 2, zl_ld(pel(r, j)));
        }
        // (d) the panel's row swaps on the columns right of it and on B,
        // (e) U12 = L11^-1 A12 on the same columns: one thread per column
        for (int c = kb + w + t; c < ncol; c += ZL_BS) {
            for (int j = 0; j < w; j++) {
                const int p = piv[j];
                if (p != j) {
                    double* x = elem(kb + j, c);
                    double* y = elem(kb + p, c);
                    const zc xv = zl_ld(x), yv = zl_ld(y);
                    zl_st(x, yv);
                    zl_st(y, xv);
                }
            }
            zc x[ZB];
#pragma unroll
            for (int j = 0; j < ZB; j++)
                if (j < w) x[j] = zl_ld(elem(kb + j, c));
#pragma unroll
            for (int j = 0; j < ZB; j++)
#pragma unroll
                for (int i = j + 1; i < ZB; i++)
                    if (i < w) x[i] = zl_msub(x[i], zl_ld(pel(i, j)), x[j]);
#pragma unroll
            for (int j = 0; j < ZB; j++)
                if (j < w) zl_st(elem(kb + j, c), x[j]);
        }
        __syncthreads();
        // (f) trailing block: A22 -= L21 U12, wavefront wv takes rows
        // kb + w + wv, + 4, ...; lane = column within a 64-column tile.
        // ZL_ILP rows at a time, each with its own accumulator: the 16-term
        // sums are chains of dependent multiply-adds, and several in flight
        // hide their latency (each sum keeps its order)
        const int c0 = kb + w;
        for (int tc = c0; tc < ncol; tc += 64) {
            const int c = tc + lane;
            if (c >= ncol) continue;  // no barrier below in this loop
            zc u12[ZB];
#pragma unroll
            for (int j = 0; j < ZB; j++) u12[j] = (j < w) ? zl_ld(elem(kb + j, c)) : zc{0.0, 0.0};
            for (int r = w + wv; r < m; r += 4 * ZL_ILP) {
                int rr[ZL_ILP];
                zc acc[ZL_ILP];
#pragma unroll
                for (int u = 0; u < ZL_ILP; u++) {
                    rr[u] = r + 4 * u;
                    acc[u] = rr[u] < m ? zl_ld(elem(kb + rr[u], c)) : zc{0.0, 0.0};
                    if (rr[u] >= m) rr[u] = m - 1;  // a valid panel row; its result is not stored
                }
#pragma unroll
                for (int j = 0; j < ZB; j++)
                    if (j < w)
#pragma unroll
                        for (int u = 0; u < ZL_ILP; u++) acc[u] = zl_msub(acc[u], zl_ld(pel(rr[u], j)), u12[j]);
#pragma unroll
                for (int u = 0; u < ZL_ILP; u++)
                    if (r + 4 * u < m) zl_st(elem(kb + r + 4 * u, c), acc[u]);
            }
        }
        __syncthreads();
    }
    // back substitution by blocks of ZB rows, last block first
    const int nblk = (n + ZB - 1) / ZB;
    for (int bk = nblk - 1; bk >= 0; bk--) {
        const int kb = bk * ZB, w = min(ZB, n - kb);
        // U block columns kb .. kb+w-1 for every row < kb + w to LDS
        for (int q = t; q < (kb + w) * w; q += ZL_BS) {
            const int r = q / w, j = q - (q / w) * w;
            zl_st(pel(r, j), zl_ld(a + ((long)r * n + kb + j) * 2));
        }
        __syncthreads();
        // the diagonal block: X rows kb .. kb+w-1, one thread per right-hand side
        for (int rr = t; rr < nrhs; rr += ZL_BS) {
            zc x[ZB];
#pragma unroll
            for (int j = 0; j < ZB; j++)
                if (j < w) x[j] = zl_ld(b + ((long)(kb + j) * nrhs + rr) * 2);
#pragma unroll
            for (int j = ZB - 1; j >= 0; j--) {
                if (j >= w) continue;
                const zc u = zl_ld(pel(kb + j, j));
                const zc ui = (u.r == 0.0 && u.i == 0.0) ? zc{0.0, 0.0} : zl_inv(u);
                x[j] = zl_mul(x[j], ui);
#pragma unroll
                for (int i = 0; i < j; i++) x[i] = zl_msub(x[i], zl_ld(pel(kb + i, j)), x[j]);
            }
#pragma unroll
            for (int j = 0; j < ZB; j++)
                if (j < w) {
                    zl_st(b + ((long)(kb + j) * nrhs + rr) * 2, x[j]);
                    zl_st(XB + ((long)j * nrhs + rr) * 2, x[j]);
                }
        }
        __syncthreads();
        // rows above: B[i][r] -= sum_j U[i][kb+j] X[kb+j][r], items (i, r) with r
        // fastest, four items per thread in flight (each sum keeps its order)
        const long nit = (long)kb * nrhs;
        for (long q0 = t; q0 < nit; q0 += 4 * ZL_BS) {
            int iu[4], ru[4];
            bool vu[4];
            zc acc[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                long q = q0 + (long)u * ZL_BS;
                vu[u] = q < nit;
                if (!vu[u]) q = nit - 1;  // a valid item; its result is not stored
                iu[u] = (int)(q / nrhs);
                ru[u] = (int)(q - (long)iu[u] * nrhs);
                acc[u] = zl_ld(b + ((long)iu[u] * nrhs + ru[u]) * 2);
            }
#pragma unroll
            for (int j = 0; j < ZB; j++)
                if (j < w)
#pragma unroll
                    for (int u = 0; u < 4; u++)
                        acc[u] = zl_msub(acc[u], zl_ld(pel(iu[u], j)), zl_ld(XB + ((long)j * nrhs + ru[u]) * 2));
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (vu[u]) zl_st(b + ((long)iu[u] * nrhs + ru[u]) * 2, acc[u]);
        }
        __syncthreads();
    }
    if (t == 0) {
        if (info) info[sys] = inf;
        if (umin) umin[sys] = mn;
        if (umax) umax[sys] = mx;
    }
}

static long zb_lds_bytes(int n, int nrhs) { return (long)(n * ZB + ZB * nrhs) * 2 * 8; }

static long zl_lds_bytes(int n, int nrhs) { return (long)(2 * n + 2 * nrhs) * 2 * 8; }

extern "C" int mm_zgesv_batched(int n, int nrhs, long batch, double* A, double* B, int* info, double* umin,
                                double* umax, void* stream) {
    mm_clear_error();
    if (n < 1 || n > ZL_MAX_N || nrhs < 1 || nrhs > ZL_MAX_N || batch < 0)
        return mm_fail(MM_EINVAL, "zgesv_batched: 1 <= n, nrhs <= %d and batch >= 0 required", ZL_MAX_N);
    if (batch == 0) return MM_OK;
    if (!A || !B) return mm_fail(MM_EINVAL, "zgesv_batched: null pointer");
    if (batch > 0x7fffffffl) return mm_fail(MM_EINVAL, "zgesv_batched: batch too large");
    if (n <= ZB_MAX_N && nrhs <= ZB_MAX_N) {
        const long lds = zb_lds_bytes(n, nrhs);
        if (lds > 65536 && hipFuncSetAttribute((const void*)k_zgesv_blk, hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)lds) != hipSuccess)
            return mm_fail(MM_EHIP, "zgesv_batched: %ld bytes of LDS for n = %d, nrhs = %d not available", lds, n,
                           nrhs);
        hipLaunchKernelGGL(k_zgesv_blk, dim3((unsigned)batch), dim3(ZL_BS), (size_t)lds, (hipStream_t)stream, n,
                           nrhs, A, B, info, umin, umax);
        return mm_check_launch("k_zgesv_blk");
    }
    const long lds = zl_lds_bytes(n, nrhs);
    if (lds > 65536 &&
        hipFuncSetAttribute((const void*)k_zgesv, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return mm_fail(MM_EHIP, "zgesv_batched: %ld bytes of LDS for n = %d, nrhs = %d not available", lds, n, nrhs);
    hipLaunchKernelGGL(k_zgesv, dim3((unsigned)batch), dim3(ZL_BS), (size_t)lds, (hipStream_t)stream, n, nrhs, A, B,
                       info, umin, umax);
    return mm_check_launch("k_zgesv");
}
