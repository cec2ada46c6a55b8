// Sine and cosine for the fast Kerr integrator, written from the
// mathematics: Cody and Waite's additive range reduction by multiples of
// pi/2 (W. J. Cody and W. Waite, Software Manual for the Elementary
// Functions, 1980, ch. 8) with the reduction done in fused multiply-adds,
// then the Taylor series of sin and cos on [-pi/4, pi/4] truncated where the
// next term is below 2^-28 (the sin series to r^9, the cos series to r^10),
// evaluated by Horner's rule. Every operation is an explicit fmaf or a single
// rounded operation, so host and device builds round identically, which the
// CPU test of this function relies on.
//
// Valid for |x| < 2^12 (the march's theta stays near [0, pi] and its phi
// within a few hundred radians); the error is a few ulp there.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define GARG_HD __host__ __device__ __forceinline__
#else
#define GARG_HD static inline
#endif

namespace garg {

// pi/2 as a sum of three floats, each the float nearest to what the
// previous ones leave (A + B + C = pi/2 to about 2^-74). Each reduction step
// is one fused multiply-add, so k * A, k * B and k * C enter exactly.
#define GARG_PIO2_A 1.5707963705062866f
#define GARG_PIO2_B -4.3711388286737929e-08f
#define GARG_PIO2_C -1.7151245100058819e-15f
#define GARG_TWO_OVER_PI 0.636619772367581343f

GARG_HD void sincos_reduced(float r, float* s, float* c) {
    float r2 = r * r;
    // sin r = r (1 - r^2/3! + r^4/5! - r^6/7! + r^8/9!)
    float ps = fmaf(r2, 2.75573192239858907e-06f, -1.98412698412698413e-04f);
    ps = fmaf(r2, ps, 8.33333333333333333e-03f);
    ps = fmaf(r2, ps, -1.66666666666666667e-01f);
    ps = ps * r2;
    *s = fmaf(r, ps, r);
    // cos r = 1 - r^2/2! + r^4/4! - r^6/6! + r^8/8! - r^10/10!
    float pc = fmaf(r2, -2.75573192239858907e-07f, 2.48015873015873016e-05f);
    pc = fmaf(r2, pc, -1.38888888888888889e-03f);
    pc = fmaf(r2, pc, 4.16666666666666667e-02f);
    pc = fmaf(r2, pc, -0.5f);
    *c = fmaf(r2, pc, 1.0f);
}

GARG_HD void fast_sincos(float x, float* s, float* c) {
    float k = rintf(x * GARG_TWO_OVER_PI);
    float r = fmaf(-k, GARG_PIO2_A, x);
    r = fmaf(-k, GARG_PIO2_B, r);
    r = fmaf(-k, GARG_PIO2_C, r);
    float sr, cr;
    sincos_reduced(r, &sr, &cr);
    int q = (int)k & 3;
    float ss = (q & 1) ? cr : sr;
    float cc = (q & 1) ? sr : cr;
    *s = (q & 2) ? -ss : ss;
    *c = ((q + 1) & 2) ? -cc : cc;
}

// sin and cos of x0 + d from s0 = sin x0 and c0 = cos x0 by the angle
// addition formulas, with sin d and cos d from their Taylor series (to d^7
// and d^6; the first omitted terms are below 1.1e-11 and 4e-10 for
// |d| <= 1/4). Shifts beyond 1/4 fall back to fast_sincos(x), where x is
// the caller's x0 + d. The Runge-Kutta stages of the fast integrator move
// theta by such small shifts from a point whose sine and cosine are known.
#define GARG_SHIFT_LIMIT 0.25f

GARG_HD void sincos_shift(float s0, float c0, float d, float x, float* s, float* c) {
    if (fabsf(d) <= GARG_SHIFT_LIMIT) {
        float d2 = d * d;
        float ps = fmaf(d2, -1.98412698412698413e-04f, 8.33333333333333333e-03f);
        ps = fmaf(d2, ps, -1.66666666666666667e-01f);
        ps = ps * d2;
        float sd = fmaf(d, ps, d);
        float pc = fmaf(d2, -1.38888888888888889e-03f, 4.16666666666666667e-02f);
        pc = fmaf(d2, pc, -0.5f);
        float cd = fmaf(d2, pc, 1.0f);
        float cs = c0 * sd;
        float ss = s0 * sd;
        *s = fmaf(s0, cd, cs);
        *c = fmaf(c0, cd, -ss);
    } else {
        fast_sincos(x, s, c);
    }
}

// 1 - cos d: its Taylor series to d^8 for |d| <= 1/4 (the first omitted
// term is below 9e-12 of the result), otherwise from fast_sincos, where the
// subtraction loses nothing that matters. The series keeps full relative
// accuracy for tiny d, where 1 - cos d in float would round to zero.
GARG_HD float versine(float d) {
    if (fabsf(d) <= GARG_SHIFT_LIMIT) {
        float d2 = d * d;
        float p = fmaf(d2, -2.48015873015873016e-05f, 1.38888888888888889e-03f);
        p = fmaf(d2, p, -4.16666666666666667e-02f);
        p = fmaf(d2, p, 0.5f);
        return d2 * p;
    }
    float s, c;
    fast_sincos(d, &s, &c);
    return 1.0f - c;
}

// atan x for the small arguments of the Kerr-Schild frame rotation: its
// Taylor series to x^9 for |x| <= 1/8 (the first omitted term, x^11 / 11,
// is below 1e-10 of the result), otherwise atanf.
GARG_HD float small_atan(float x) {
    if (fabsf(x) <= 0.125f) {
        float x2 = x * x;
        float p = fmaf(x2, 1.11111111111111111e-01f, -1.42857142857142857e-01f);
        p = fmaf(x2, p, 2.0e-01f);
        p = fmaf(x2, p, -3.33333333333333333e-01f);
        p = p * x2;
        return fmaf(x, p, x);
    }
    return atanf(x);
}

// ln |1 + e| for the small ratios of the Kerr-Schild frame rotation: its
// series to e^8 for |e| <= 1/16 (the first omitted term, e^9 / 9, is below
// 1e-10 of the result), otherwise log1pf (logf of |1 + e| where 1 + e
// falls below 1/2).
GARG_HD float small_log1p(float e) {
    if (fabsf(e) <= 0.0625f) {
        float p = fmaf(e, -0.125f, 1.42857142857142857e-01f);
        p = fmaf(e, p, -1.66666666666666667e-01f);
        p = fmaf(e, p, 2.0e-01f);
        p = fmaf(e, p, -0.25f);
        p = fmaf(e, p, 3.33333333333333333e-01f);
        p = fmaf(e, p, -0.5f);
        p = p * e;
        return fmaf(e, p, e);
    }
    return e > -0.5f ? log1pf(e) : logf(fmaxf(fabsf(1.0f + e), 1e-30f));
}

// sin d: its Taylor series to d^9 for |d| <= 1/4 (the first omitted term is
// below 4e-13 of the result), otherwise from fast_sincos.
GARG_HD float small_sin(float d) {
    if (fabsf(d) <= GARG_SHIFT_LIMIT) {
        float d2 = d * d;
        float p = fmaf(d2, 2.75573192239858907e-06f, -1.98412698412698413e-04f);
        p = fmaf(d2, p, 8.33333333333333333e-03f);
        p = fmaf(d2, p, -1.66666666666666667e-01f);
        p = p * d2;
        return fmaf(d, p, d);
    }
    float s, c;
    fast_sincos(d, &s, &c);
    return s;
}

}  // namespace garg
