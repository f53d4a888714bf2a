// float32 sin/cos, exp and log for the batched fast path, bit-reproducible on
// the host compiler: Cody-Waite reduction (pi/2 in three parts, ln 2 in two)
// with fused multiply-adds, then Taylor polynomials of sin, cos (to x^9,
// x^10) and exp (to x^7) in Horner form; log splits x = 2^e m with
// m in [sqrt(1/2), sqrt(2)) and sums the atanh series of s = (m-1)/(m+1) to
// s^11.  Errors: sin/cos <= 1.5 ulp on [-2 pi, 2 pi], exp <= 1 ulp for
// normal results, log <= 1.5 ulp for normal x.
#pragma once

#include <hip/hip_runtime.h>

namespace mc {

__device__ __forceinline__ void fast_sincosf(float x, float &sn, float &cs) {
    const float k = rintf(x * 0x1.45f306p-1f);
    float r = fmaf(-k, 0x1.921fb6p+0f, x);
    r = fmaf(-k, -0x1.777a5cp-25f, r);
    r = fmaf(-k, -0x1p-49f, r);
    const float r2 = r * r;
    float ps = fmaf(r2, 0x1.71de3ap-19f, -0x1.a01a02p-13f);
    ps = fmaf(r2, ps, 0x1.111112p-7f);
    ps = fmaf(r2, ps, -0x1.555556p-3f);
    const float s = fmaf(r * r2, ps, r);
    float pc = fmaf(r2, -0x1.27e4fcp-22f, 0x1.a01a02p-16f);
    pc = fmaf(r2, pc, -0x1.6c16c2p-10f);
    pc = fmaf(r2, pc, 0x1.555556p-5f);
    pc = fmaf(r2, pc, -0x1p-1f);
    const float c = fmaf(r2, pc, 1.0f);
    const int q = ((int)k) & 3;
    const float s1 = (q & 1) ? c : s;
    const float c1 = (q & 1) ? s : c;
    sn = (q & 2) ? -s1 : s1;
    cs = ((q + 1) & 2) ? -c1 : c1;
}

// Branch-free: the arguments at or below -104 (and NaN) take the select at
// the end, so an inlined call adds no control flow to the caller.
__device__ __forceinline__ float fast_expf(float x) {
    const bool tiny = !(x > -104.0f);  // also NaN
    const float xr = tiny ? 0.0f : x;
    const float k = rintf(xr * 0x1.715476p+0f);
    float r = fmaf(-k, 0x1.62e430p-1f, xr);
    r = fmaf(-k, -0x1.05c610p-29f, r);
    float p = fmaf(r, 0x1.a01a02p-13f, 0x1.6c16c2p-10f);
    p = fmaf(p, r, 0x1.111112p-7f);
    p = fmaf(p, r, 0x1.555556p-5f);
    p = fmaf(p, r, 0x1.555556p-3f);
    p = fmaf(p, r, 0x1p-1f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int e = (int)k;
    const bool low = e < -125;
    p = low ? p * 0x1p-64f : p;
    e = low ? e + 64 : e;
    const float v = p * __int_as_float((e + 127) << 23);
    return tiny ? (x != x ? x : 0.0f) : v;
}

// Metropolis acceptance weight exp(min(x, 0)) for the fast path, branch free:
// x is clamped to [-87, 0], where every result is a normal float (exp(0) = 1
// exactly, so "dE <= 0 or v < exp(-beta dE)" is "v < accept_weight(-beta dE)"
// for any v in [0, 1)).  Same reduction and polynomial as fast_expf.
__device__ __forceinline__ float accept_weight(float x) {
    x = fminf(fmaxf(x, -87.0f), 0.0f);
    const float k = rintf(x * 0x1.715476p+0f);
    float r = fmaf(-k, 0x1.62e430p-1f, x);
    r = fmaf(-k, -0x1.05c610p-29f, r);
    float p = fmaf(r, 0x1.a01a02p-13f, 0x1.6c16c2p-10f);
    p = fmaf(p, r, 0x1.111112p-7f);
    p = fmaf(p, r, 0x1.555556p-5f);
    p = fmaf(p, r, 0x1.555556p-3f);
    p = fmaf(p, r, 0x1p-1f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return p * __int_as_float(((int)k + 127) << 23);
}

// log for positive normal x (the engine calls it on uniforms in (0, 1])
__device__ __forceinline__ float fast_logf(float x) {
    const unsigned b = __float_as_uint(x);
    int e = (int)(b >> 23) - 127;
    float m = __uint_as_float((b & 0x7fffffu) | 0x3f800000u);  // [1, 2)
    if (m > 0x1.6a09e6p+0f) {                                  // sqrt(2)
        m *= 0.5f;
        e += 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float p = fmaf(z, 0x1.745d18p-3f, 0x1.c71c72p-3f);  // 2/11, 2/9
    p = fmaf(z, p, 0x1.24924ap-2f);                     // 2/7
    p = fmaf(z, p, 0x1.99999ap-2f);                     // 2/5
    p = fmaf(z, p, 0x1.555556p-1f);                     // 2/3
    const float r = z * p;
    const float lm = fmaf(-s, f - r, f);  // 2 atanh(s) = f - s (f - r)
    const float ef = (float)e;
    return fmaf(ef, 0x1.62e430p-1f, fmaf(ef, -0x1.05c610p-29f, lm));
}

// sqrt(y) for y >= 0 as y * rsqrt(y): seed from the exponent/mantissa bits
// (Lomont's constant 0x5f375a86), three Newton steps x <- x (1.5 - y/2 x^2) with
// fused multiply-adds.  Relative error < 2e-7; exactly reproducible on a CPU.
__device__ __forceinline__ float newton_sqrtf(float y) {
    float x = __uint_as_float(0x5f375a86u - (__float_as_uint(y) >> 1));
    const float h = 0.5f * y;
    x = x * fmaf(-(h * x), x, 1.5f);
    x = x * fmaf(-(h * x), x, 1.5f);
    x = x * fmaf(-(h * x), x, 1.5f);
    return y * x;
}

// sin(x) for |x| <= pi/2: odd Taylor series to x^13 in Horner form (error < 7e-10).
__device__ __forceinline__ float halfpi_sinf(float x) {
    const float x2 = x * x;
    float p = fmaf(x2, 0x1.612462p-33f, -0x1.ae6456p-26f);  // 1/13!, -1/11!
    p = fmaf(x2, p, 0x1.71de3ap-19f);                        // 1/9!
    p = fmaf(x2, p, -0x1.a01a02p-13f);                       // -1/7!
    p = fmaf(x2, p, 0x1.111112p-7f);                         // 1/5!
    p = fmaf(x2, p, -0x1.555556p-3f);                        // -1/3!
    return fmaf(x * x2, p, x);
}

// Gaussian angle step of the batched fast path from ONE 64-bit draw x:
//   bits 63..43 -> k1, u1 = (k1 + 1) 2^-21 in (0, 1]
//   bits 42..22 -> sign bit b and 20-bit a: phi = pi (a 2^-20 - 1/2) in [-pi/2, pi/2)
//   bits 21..1  -> k3, acceptance uniform v = k3 2^-21 in [0, 1)
// z = (-1)^b sqrt(-2 log u1) sin(phi): Box-Muller with the sine's sign taken
// from an independent bit, so the proposal is exactly symmetric.
struct GaussDraw {
    float z, v;
};
__device__ __forceinline__ GaussDraw gauss_draw(uint64_t x) {
    const unsigned hi = (unsigned)(x >> 32), lo = (unsigned)x;
    const unsigned k1 = hi >> 11;
    const unsigned k2 = ((hi & 0x7ffu) << 10) | (lo >> 22);
    const unsigned k3 = (lo >> 1) & 0x1fffffu;
    const float u1 = (float)(k1 + 1u) * 0x1p-21f;
    const float phi = fmaf((float)(k2 & 0xfffffu), 0x1.921fb6p-19f, -0x1.921fb6p+0f);  // pi 2^-20, -pi/2
    const float z = newton_sqrtf(-2.0f * fast_logf(u1)) * halfpi_sinf(phi);
    GaussDraw d;
    d.z = (k2 >> 20) ? -z : z;
    d.v = (float)k3 * 0x1p-21f;
    return d;
}

// Angle into [-pi, pi) after an unbounded move: whole turns, then the last
// one's rounding (branch-free).
__device__ __forceinline__ float wrap_turns(float p) {
    const float pi = 3.14159274101257324f, two_pi = 6.28318548202514648f;
    p -= two_pi * floorf((p + pi) * 0x1.45f306p-3f);  // 1 / (2 pi)
    p = p >= pi ? p - two_pi : p;
    return p < -pi ? p + two_pi : p;
}

}  // namespace mc
