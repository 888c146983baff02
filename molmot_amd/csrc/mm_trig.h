// (cos, sin) helpers shared by the MCWF and Lindblad kernels.
#pragma once

#include <hip/hip_runtime.h>

// (cos, sin) of theta for amplitude arithmetic in FP32: FP64 reduction to
// revolutions, then the hardware sin/cos (argument in revolutions).
__device__ __forceinline__ void mm_cis(double theta, float& c, float& s) {
    double x = theta * 0.15915494309189535;  // revolutions
    x -= rint(x);
    const float xf = (float)x;
    c = __builtin_amdgcn_cosf(xf);
    s = __builtin_amdgcn_sinf(xf);
}
// FP64 (cos, sin) for |theta| < 2^20: Cody-Waite reduction by pi/2 with the
// two-part split of fdlibm (33 + 33 bits, so k * PIO2_1 is exact), then the
// Taylor series of sin and cos on |x| <= pi/4 (truncation < 2e-19) and a
// quadrant swap.
__device__ __forceinline__ void mm_cis(double theta, double& c, double& s) {
    const double PIO2_1 = 1.57079632673412561417e+00, PIO2_1T = 6.07710050650619224932e-11;
    const double k = rint(theta * 0.63661977236758134308);
    const double x = fma(-k, PIO2_1T, fma(-k, PIO2_1, theta));
    const double z = x * x;
    // sin x = x (1 - z/3! + z^2/5! - ... - z^8/17!), cos x = 1 - z/2! + ... + z^9/18!
    double ps = -1.0 / 355687428096000.0;
    ps = fma(ps, z, 1.0 / 1307674368000.0);
    ps = fma(ps, z, -1.0 / 6227020800.0);
    ps = fma(ps, z, 1.0 / 39916800.0);
    ps = fma(ps, z, -1.0 / 362880.0);
    ps = fma(ps, z, 1.0 / 5040.0);
    ps = fma(ps, z, -1.0 / 120.0);
    ps = fma(ps, z, 1.0 / 6.0);
    const double sx = fma(-x * z, ps, x);
    double pc = 1.0 / 6402373705728000.0;
    pc = fma(pc, z, -1.0 / 20922789888000.0);
    pc = fma(pc, z, 1.0 / 87178291200.0);
    pc = fma(pc, z, -1.0 / 479001600.0);
    pc = fma(pc, z, 1.0 / 3628800.0);
    pc = fma(pc, z, -1.0 / 40320.0);
    pc = fma(pc, z, 1.0 / 720.0);
    pc = fma(pc, z, -1.0 / 24.0);
    pc = fma(pc, z, 0.5);
    const double cx = fma(-z, pc, 1.0);
    const int q = ((int)k) & 3;
    const double sa = (q & 1) ? cx : sx, ca = (q & 1) ? sx : cx;
    s = (q & 2) ? -sa : sa;
    c = ((q + 1) & 2) ? -ca : ca;
}

