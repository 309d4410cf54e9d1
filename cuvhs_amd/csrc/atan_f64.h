// FP64 two-argument arctangent for K1's instantaneous phase.
//
// Method: with a = min(|y|, |x|) and b = max, atan(a/b) is evaluated on
// |t| <= tan(pi/8) by the identity atan(s) = pi/4 + atan((s - 1)/(s + 1)),
// folded into the one division (t = a/b, or t = (a - b)/(a + b) when
// a > tan(pi/8) b), then atan(t) = t + t u P(u), u = t^2, with P the
// degree-10 Chebyshev interpolant of (atan(sqrt u)/sqrt u - 1)/u on
// [0, tan^2(pi/8)] (coefficients from gen/atan_coeffs.py; interpolation
// error below 1e-17 relative to atan(t)).  Octant and sign: pi/2 - r when
// |y| > |x|, pi - r for x with its sign bit set, the sign of y last, which
// gives the IEEE results for signed zeros (atan2(+-0, +0) = +-0,
// atan2(+-0, -0) = +-pi).  Inputs here are finite.  Error: a few ulp of the
// result (the division, the two reduction roundings and the final
// additions), checked against the host's atan2 by the GPU tests.
//
// The polynomial's coefficients are uniform, so each Horner step is one
// v_fma_f64 with the coefficient in an SGPR pair (scalar moves set it);
// written as fma() the compiler would copy each one into VGPRs for the
// two-operand v_fmac form, three vector instructions per term instead of one.
#pragma once
#include <hip/hip_runtime.h>

namespace cuvhs {

__device__ __forceinline__ double fma_sgpr(double a, double b, double c) {
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(c));
    return r;
}

__device__ __forceinline__ double atan2_f64(double y, double x) {
    const double ax = fabs(x), ay = fabs(y);
    const bool sw = ay > ax;
    const double a = sw ? ax : ay, b = sw ? ay : ax;
    const bool red = a > 0.41421356237309503 * b;
    const double num = red ? a - b : a;
    double den = red ? a + b : b;
    den = den > 0.0 ? den : 1.0;  // a = b = 0: t = 0
    const double t = num / den;
    const double u = t * t;
    double p = -0.01917688711906226;
    p = fma_sgpr(p, u, 0.03923165829558719);
    p = fma_sgpr(p, u, -0.0508544973794026);
    p = fma_sgpr(p, u, 0.0585814891280221);
    p = fma_sgpr(p, u, -0.06664511447381948);
    p = fma_sgpr(p, u, 0.07692183190826087);
    p = fma_sgpr(p, u, -0.09090904578123903);
    p = fma_sgpr(p, u, 0.11111111015256361);
    p = fma_sgpr(p, u, -0.14285714284666542);
    p = fma_sgpr(p, u, 0.1999999999999552);
    p = fma_sgpr(p, u, -0.3333333333333333);
    double r = fma(t * u, p, t);
    if (red) r += 0.7853981633974483;
    if (sw) r = 1.5707963267948966 - r;
    if (signbit(x)) r = 3.141592653589793 - r;
    return copysign(r, y);
}

}  // namespace cuvhs
