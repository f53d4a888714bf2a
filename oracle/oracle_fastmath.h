/*
 * The oracle's float32 sin/cos, exp and log (test infrastructure only), as
 * inline functions for both oracle sources: Cody-Waite reduction by pi/2
 * (three-part split) or ln 2 (two-part split) with fused multiply-adds, then
 * the Taylor series of sin, cos (to x^9, x^10) and exp (to x^7) in Horner
 * form; log splits x = 2^e m, m in [sqrt(1/2), sqrt(2)), and sums the atanh
 * series of s = (m-1)/(m+1).  The HIP kernels evaluate the same operations
 * in the same order, so both produce identical bits.
 */
#ifndef ORACLE_FASTMATH_H
#define ORACLE_FASTMATH_H
#include <math.h>
#include <stdint.h>

static const float PIO2_A = 0x1.921fb6p+0f;
static const float PIO2_B = -0x1.777a5cp-25f;
static const float PIO2_C = -0x1p-49f;
static const float TWO_OVER_PI = 0x1.45f306p-1f;
static const float S1 = -0x1.555556p-3f, S2 = 0x1.111112p-7f, S3 = -0x1.a01a02p-13f, S4 = 0x1.71de3ap-19f;
static const float C1 = -0x1p-1f, C2 = 0x1.555556p-5f, C3 = -0x1.6c16c2p-10f, C4 = 0x1.a01a02p-16f,
                   C5 = -0x1.27e4fcp-22f;
static const float LOG2E = 0x1.715476p+0f;
static const float LN2_HI = 0x1.62e430p-1f, LN2_LO = -0x1.05c610p-29f;
static const float E2 = 0x1p-1f, E3 = 0x1.555556p-3f, E4 = 0x1.555556p-5f, E5 = 0x1.111112p-7f,
                   E6 = 0x1.6c16c2p-10f, E7 = 0x1.a01a02p-13f;

/* valid for |x| <= 2^20; the engine only calls it on angles in [-2 pi, 2 pi] */
static inline void ofm_sincosf(float x, float *sn, float *cs) {
    const float k = rintf(x * TWO_OVER_PI);
    float r = fmaf(-k, PIO2_A, x);
    r = fmaf(-k, PIO2_B, r);
    r = fmaf(-k, PIO2_C, r);
    const float r2 = r * r;
    float ps = fmaf(r2, S4, S3);
    ps = fmaf(r2, ps, S2);
    ps = fmaf(r2, ps, S1);
    const float s = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, C5, C4);
    pc = fmaf(r2, pc, C3);
    pc = fmaf(r2, pc, C2);
    pc = fmaf(r2, pc, C1);
    const float c = fmaf(r2, pc, 1.0f);
    const int q = ((int)k) & 3;
    if (q == 0) { *sn = s; *cs = c; }
    else if (q == 1) { *sn = c; *cs = -s; }
    else if (q == 2) { *sn = -s; *cs = -c; }
    else { *sn = -c; *cs = s; }
}

/* exp for x <= 88; NaN propagates, x below -104 gives 0 */
static inline float ofm_expf(float x) {
    if (!(x > -104.0f)) return x != x ? x : 0.0f;
    const float k = rintf(x * LOG2E);
    float r = fmaf(-k, LN2_HI, x);
    r = fmaf(-k, LN2_LO, r);
    float p = fmaf(r, E7, E6);
    p = fmaf(p, r, E5);
    p = fmaf(p, r, E4);
    p = fmaf(p, r, E3);
    p = fmaf(p, r, E2);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int e = (int)k;
    if (e < -125) { /* two exact power-of-two scalings keep 2^e normal */
        p *= 0x1p-64f;
        e += 64;
    }
    union { uint32_t u; float f; } two_e;
    two_e.u = (uint32_t)(e + 127) << 23;
    return p * two_e.f;
}

/* log for positive normal x: x = 2^e m, m in [sqrt(1/2), sqrt(2)), then
 * log m = 2 atanh(s), s = (m-1)/(m+1), summed to s^11 as f - s (f - r) */
static inline float ofm_logf(float x) {
    union { float f; uint32_t u; } v;
    v.f = x;
    int e = (int)(v.u >> 23) - 127;
    v.u = (v.u & 0x7fffffu) | 0x3f800000u;
    float m = v.f;
    if (m > 0x1.6a09e6p+0f) {
        m *= 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float p = fmaf(z, 0x1.745d18p-3f, 0x1.c71c72p-3f);
    p = fmaf(z, p, 0x1.24924ap-2f);
    p = fmaf(z, p, 0x1.99999ap-2f);
    p = fmaf(z, p, 0x1.555556p-1f);
    const float r = z * p;
    const float lm = fmaf(-s, f - r, f);
    const float ef = (float)e;
    return fmaf(ef, LN2_HI, fmaf(ef, LN2_LO, lm));
}

#endif
