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

/* ln(n 2^-21) for an integer n in [1, 2^21] without a division (the Gaussian
 * step's radius): n = 2^E m, m in [1, 2); cell i = round(64 (m - 1)) of 65,
 * c = 1 + i/64; the table holds rc = float(1/c) and -ln(rc) = hi + lo, and
 *   ln(n 2^-21) = ((E-21) LN2_HI + hi) + (((E-21) LN2_LO + lo) + log1p(t)),
 *   t = fma(m, rc, -1), log1p(t) = t - t^2/2 + t^3/3 - t^4/4 in Horner form.
 * Cell 64 (c = 2) holds ln 2 as (LN2_HI, LN2_LO). */
static const float OFM_LN_U21[65][3] = {
    {0x1p+0f, 0x0p+0f, 0x0p+0f},
    {0x1.f81f82p-1f, 0x1.fc0a8ap-7f, -0x1.e07f84p-32f},
    {0x1.f07c2p-1f, 0x1.f8299p-6f, 0x1.cf067p-31f},
    {0x1.e9131ap-1f, 0x1.77459cp-5f, -0x1.cd22dcp-33f},
    {0x1.e1e1e2p-1f, 0x1.f0a30ap-5f, 0x1.162a76p-37f},
    {0x1.dae608p-1f, 0x1.341d74p-4f, 0x1.86f478p-30f},
    {0x1.d41d42p-1f, 0x1.6f0d28p-4f, -0x1.a35296p-29f},
    {0x1.cd8568p-1f, 0x1.a926d8p-4f, 0x1.495aaep-29f},
    {0x1.c71c72p-1f, 0x1.e27074p-4f, 0x1.c55e5cp-29f},
    {0x1.c0e07p-1f, 0x1.0d77e8p-3f, 0x1.9a11ccp-28f},
    {0x1.bacf92p-1f, 0x1.29552cp-3f, 0x1.07fd4cp-29f},
    {0x1.b4e81cp-1f, 0x1.44d2b4p-3f, -0x1.cd20b6p-29f},
    {0x1.af286cp-1f, 0x1.5ff306p-3f, 0x1.4f27aap-32f},
    {0x1.a98ef6p-1f, 0x1.7ab89p-3f, 0x1.043642p-29f},
    {0x1.a41a42p-1f, 0x1.9525a8p-3f, 0x1.e8ad7p-32f},
    {0x1.9ec8eap-1f, 0x1.af3c92p-3f, -0x1.dfdp-29f},
    {0x1.99999ap-1f, 0x1.c8ff7ap-3f, 0x1.e6a68ap-29f},
    {0x1.948b1p-1f, 0x1.e27076p-3f, -0x1.d50d1ap-31f},
    {0x1.8f9c18p-1f, 0x1.fb918cp-3f, -0x1.50e0dep-30f},
    {0x1.8acb9p-1f, 0x1.0a325p-2f, 0x1.4e721ep-27f},
    {0x1.861862p-1f, 0x1.1675cap-2f, -0x1.0a8b3ep-27f},
    {0x1.818182p-1f, 0x1.22941ep-2f, 0x1.b3de5ap-28f},
    {0x1.7d05f4p-1f, 0x1.2e8e2cp-2f, -0x1.1ee2dp-30f},
    {0x1.78a4c8p-1f, 0x1.3a64c6p-2f, -0x1.a5ae86p-28f},
    {0x1.745d18p-1f, 0x1.4618bap-2f, 0x1.0e2f66p-29f},
    {0x1.702e06p-1f, 0x1.51aad8p-2f, -0x1.e903eap-29f},
    {0x1.6c16c2p-1f, 0x1.5d1bdap-2f, 0x1.560274p-28f},
    {0x1.681682p-1f, 0x1.686c8p-2f, 0x1.cd8a5ap-29f},
    {0x1.642c86p-1f, 0x1.739d7ep-2f, 0x1.5de804p-29f},
    {0x1.605816p-1f, 0x1.7eaf84p-2f, -0x1.bea81ep-29f},
    {0x1.5c9882p-1f, 0x1.89a33ap-2f, 0x1.18284cp-27f},
    {0x1.58ed24p-1f, 0x1.94793ep-2f, 0x1.c422e2p-27f},
    {0x1.555556p-1f, 0x1.9f323cp-2f, 0x1.97f30ap-27f},
    {0x1.51d07ep-1f, 0x1.a9ceccp-2f, -0x1.197decp-28f},
    {0x1.4e5e0ap-1f, 0x1.b44f7ap-2f, -0x1.c66e14p-27f},
    {0x1.4afd6ap-1f, 0x1.beb4dap-2f, -0x1.58e484p-30f},
    {0x1.47ae14p-1f, 0x1.c8ff7ep-2f, -0x1.959768p-32f},
    {0x1.446f86p-1f, 0x1.d32fe8p-2f, 0x1.e01d7ap-27f},
    {0x1.414142p-1f, 0x1.dd469ep-2f, -0x1.3e3b54p-30f},
    {0x1.3e22ccp-1f, 0x1.e74426p-2f, -0x1.052f0ep-27f},
    {0x1.3b13b2p-1f, 0x1.f128f4p-2f, -0x1.0a1f2p-27f},
    {0x1.381382p-1f, 0x1.faf586p-2f, 0x1.9e3ccap-28f},
    {0x1.3521dp-1f, 0x1.02552ap-1f, -0x1.2d178p-28f},
    {0x1.323e34p-1f, 0x1.0723e6p-1f, 0x1.a39be8p-26f},
    {0x1.2f684cp-1f, 0x1.0be72ep-1f, 0x1.295418p-32f},
    {0x1.2c9fb4p-1f, 0x1.109f3cp-1f, -0x1.5a566cp-26f},
    {0x1.29e412p-1f, 0x1.154c3ep-1f, 0x1.fa6af8p-28f},
    {0x1.27350cp-1f, 0x1.19ee6ap-1f, 0x1.d9f25cp-27f},
    {0x1.24924ap-1f, 0x1.1e85f4p-1f, 0x1.9c1036p-27f},
    {0x1.21fb78p-1f, 0x1.23130ep-1f, -0x1.90503p-27f},
    {0x1.1f7048p-1f, 0x1.2795ep-1f, 0x1.d13624p-26f},
    {0x1.1cf06ap-1f, 0x1.2c0eap-1f, 0x1.7123a6p-27f},
    {0x1.1a7b96p-1f, 0x1.307d74p-1f, -0x1.561de8p-26f},
    {0x1.181182p-1f, 0x1.34e288p-1f, 0x1.8e70fp-28f},
    {0x1.15b1e6p-1f, 0x1.393e0ep-1f, -0x1.b53abcp-26f},
    {0x1.135c82p-1f, 0x1.3d9024p-1f, 0x1.de2aep-26f},
    {0x1.111112p-1f, 0x1.41d8fcp-1f, 0x1.88ce58p-26f},
    {0x1.0ecf56p-1f, 0x1.4618bep-1f, -0x1.d8e84ep-27f},
    {0x1.0c9714p-1f, 0x1.4a4f88p-1f, -0x1.13f04ep-27f},
    {0x1.0a681p-1f, 0x1.4e7d82p-1f, 0x1.6dd6eep-27f},
    {0x1.08421p-1f, 0x1.52a2d4p-1f, -0x1.34874ap-26f},
    {0x1.0624dep-1f, 0x1.56bf9cp-1f, -0x1.e6062ep-28f},
    {0x1.041042p-1f, 0x1.5ad402p-1f, 0x1.a6b3e8p-26f},
    {0x1.020408p-1f, 0x1.5ee02ap-1f, 0x1.6482cep-26f},
    {0x1p-1f, 0x1.62e43p-1f, -0x1.05c61p-29f},
};

static inline float ofm_ln_u21(uint32_t n) {
    union { float f; uint32_t u; } v;
    v.f = (float)n;
    const uint32_t b = v.u;
    const uint32_t i = (((b >> 16) & 0x7fu) + 1u) >> 1;
    v.u = (b & 0x7fffffu) | 0x3f800000u;
    const float m = v.f;
    const float ef = (float)((int)(b >> 23) - 148);
    const float t = fmaf(m, OFM_LN_U21[i][0], -1.0f);
    const float t2 = t * t;
    float p = fmaf(t, -0.25f, 0x1.555556p-2f);
    p = fmaf(p, t, -0.5f);
    const float lp = fmaf(t2, p, t);
    const float hi = fmaf(ef, LN2_HI, OFM_LN_U21[i][1]);
    const float lo = fmaf(ef, LN2_LO, OFM_LN_U21[i][2]);
    return hi + (lo + lp);
}

#endif
