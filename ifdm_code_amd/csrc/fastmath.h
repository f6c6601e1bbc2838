// fastmath.h -- sin/cos of phase arguments for the kick and drift phases.
//
// exp(-i theta) is needed once per cell per kick.  The library's generic
// FP64 sincos carries a large-argument (Payne-Hanek) path that dominates the
// instruction count of the kick pass; phase arguments here are O(10) rad, so
// this file reduces with a three-term Cody-Waite split of pi/2 (each term
// applied with one fused multiply-add) and evaluates truncated Taylor series
// of sin and cos on |r| <= pi/4 in Horner form.  Truncation error: the first
// omitted sin term r^19/19! and cos term r^20/20! are below 1e-19 on that
// interval (FP64); r^11/11! and r^12/12! below 2e-9 and 1e-10 (FP32).
// Constants: pi/2 split into doubles p1 + p2 + p3 (p1 = pi/2 rounded, p2 =
// the rounded remainder, p3 the next remainder); 1/k! rounded to double.
// The reduction is done in FP64 for both precisions, so the mixed mode's
// FP32 sincos only ever sees |r| <= pi/4.
//
// The header is also compiled as plain C++ by the CPU tests.
#pragma once
#include <math.h>
#if !defined(__HIPCC__)
#ifndef __host__
#define __host__
#endif
#ifndef __device__
#define __device__
#endif
#ifndef __forceinline__
#define __forceinline__ inline
#endif
#endif

namespace ifdm {

struct PhaseRed {
    double r;  // theta - n pi/2, |r| <= pi/4 (+ rounding)
    int q;     // n mod 4
};

__host__ __device__ __forceinline__ PhaseRed reduce_pio2(double x) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    const double two_over_pi = 0.6366197723675814;
    const double p1 = 1.5707963267948966, p2 = 6.123233995736766e-17, p3 = -1.4973849048591698e-33;
    // 1.5 * 2^52: adding it rounds to the nearest integer (ties to even, as
    // rint) and leaves that integer, two's complement, in the low bits of the
    // significand for |x 2/pi| < 2^51; subtracting it again is exact.  The
    // quadrant is then read from the bits, one add instead of a conversion
    // of the double to a 64-bit integer.
    const double magic = 6755399441055744.0;
    const double t = x * two_over_pi + magic;
    const double n = t - magic;
    double r = fma(-n, p1, x);
    r = fma(-n, p2, r);
    r = fma(-n, p3, r);
    PhaseRed o;
    o.r = r;
    o.q = (int)__builtin_bit_cast(long long, t) & 3;
    return o;
}

// sin and cos of |r| <= pi/4
__host__ __device__ __forceinline__ void sincos_poly(double r, double *s, double *c) {
    const double r2 = r * r;
    double ps = -8.22063524662433e-18;                 // -1/19!
    ps = fma(ps, r2, 2.8114572543455206e-15);          // 1/17!
    ps = fma(ps, r2, -7.647163731819816e-13);          // -1/15!
    ps = fma(ps, r2, 1.6059043836821613e-10);          // 1/13!
    ps = fma(ps, r2, -2.505210838544172e-08);          // -1/11!
    ps = fma(ps, r2, 2.7557319223985893e-06);          // 1/9!
    ps = fma(ps, r2, -0.0001984126984126984);          // -1/7!
    ps = fma(ps, r2, 0.008333333333333333);            // 1/5!
    ps = fma(ps, r2, -0.16666666666666666);            // -1/3!
    *s = fma(ps * r2, r, r);
    double pc = -1.5619206968586225e-16;               // -1/18!
    pc = fma(pc, r2, 4.779477332387385e-14);           // 1/16!
    pc = fma(pc, r2, -1.1470745597729725e-11);         // -1/14!
    pc = fma(pc, r2, 2.08767569878681e-09);            // 1/12!
    pc = fma(pc, r2, -2.755731922398589e-07);          // -1/10!
    pc = fma(pc, r2, 2.48015873015873e-05);            // 1/8!
    pc = fma(pc, r2, -0.001388888888888889);           // -1/6!
    pc = fma(pc, r2, 0.041666666666666664);            // 1/4!
    pc = fma(pc, r2, -0.5);                            // -1/2!
    *c = fma(pc, r2, 1.0);
}

__host__ __device__ __forceinline__ void sincos_poly(float r, float *s, float *c) {
    const float r2 = r * r;
    float ps = 2.7557319e-06f;                         // 1/9!
    ps = fmaf(ps, r2, -1.9841270e-04f);                // -1/7!
    ps = fmaf(ps, r2, 8.3333333e-03f);                 // 1/5!
    ps = fmaf(ps, r2, -1.6666667e-01f);                // -1/3!
    *s = fmaf(ps * r2, r, r);
    float pc = -2.7557319e-07f;                        // -1/10!
    pc = fmaf(pc, r2, 2.4801587e-05f);                 // 1/8!
    pc = fmaf(pc, r2, -1.3888889e-03f);                // -1/6!
    pc = fmaf(pc, r2, 4.1666667e-02f);                 // 1/4!
    pc = fmaf(pc, r2, -0.5f);                          // -1/2!
    *c = fmaf(pc, r2, 1.0f);
}

// (cos theta, -sin theta) = exp(-i theta) in precision T; theta in FP64
template <typename T>
__host__ __device__ __forceinline__ void expmi_parts(double theta, T *re, T *im) {
    const PhaseRed red = reduce_pio2(theta);
    T s, c;
    sincos_poly((T)red.r, &s, &c);
    // theta = r + q pi/2: (cos, sin) of theta is (c, s), (-s, c), (-c, -s),
    // (s, -c) for q = 0..3; written as selects so that the lanes of a
    // wavefront, whose quadrants differ, take no branch
    const bool odd = red.q & 1;
    T ct = odd ? s : c;
    T st = odd ? c : s;
    if ((red.q + 1) & 2) ct = -ct;
    if (red.q & 2) st = -st;
    *re = ct;
    *im = -st;
}

}  // namespace ifdm
