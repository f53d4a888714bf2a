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

// ln(n 2^-21) for an integer n in [1, 2^21], without a division (the
// Gaussian step's radius).  n = 2^E m with m in [1, 2); i = round(64 (m - 1))
// picks one of 65 cells c = 1 + i/64, so |m rc - 1| <= 2^-7 with
// rc = float(1/c); the table holds rc and -ln(rc) = hi + lo (two floats), and
//   ln(n 2^-21) = ((E - 21) ln2_hi + hi) + (((E - 21) ln2_lo + lo) + log1p(t)),
//   t = m rc - 1 (one rounding), log1p(t) = t - t^2/2 + t^3/3 - t^4/4.
// Cell 64 (c = 2) holds ln 2 in the constant's own two parts, so for n just
// below 2^21 the large terms cancel exactly.  Error <= 1.3 ulp over all 2^21
// inputs.  The kernels copy the table to LDS (ln_table_to_lds) and index it
// there.
constexpr int kLnCells = 65;
static __constant__ float4 kLnU21[kLnCells] = {
    {0x1p+0f, 0x0p+0f, 0x0p+0f, 0.0f},
    {0x1.f81f82p-1f, 0x1.fc0a8ap-7f, -0x1.e07f84p-32f, 0.0f},
    {0x1.f07c2p-1f, 0x1.f8299p-6f, 0x1.cf067p-31f, 0.0f},
    {0x1.e9131ap-1f, 0x1.77459cp-5f, -0x1.cd22dcp-33f, 0.0f},
    {0x1.e1e1e2p-1f, 0x1.f0a30ap-5f, 0x1.162a76p-37f, 0.0f},
    {0x1.dae608p-1f, 0x1.341d74p-4f, 0x1.86f478p-30f, 0.0f},
    {0x1.d41d42p-1f, 0x1.6f0d28p-4f, -0x1.a35296p-29f, 0.0f},
    {0x1.cd8568p-1f, 0x1.a926d8p-4f, 0x1.495aaep-29f, 0.0f},
    {0x1.c71c72p-1f, 0x1.e27074p-4f, 0x1.c55e5cp-29f, 0.0f},
    {0x1.c0e07p-1f, 0x1.0d77e8p-3f, 0x1.9a11ccp-28f, 0.0f},
    {0x1.bacf92p-1f, 0x1.29552cp-3f, 0x1.07fd4cp-29f, 0.0f},
    {0x1.b4e81cp-1f, 0x1.44d2b4p-3f, -0x1.cd20b6p-29f, 0.0f},
    {0x1.af286cp-1f, 0x1.5ff306p-3f, 0x1.4f27aap-32f, 0.0f},
    {0x1.a98ef6p-1f, 0x1.7ab89p-3f, 0x1.043642p-29f, 0.0f},
    {0x1.a41a42p-1f, 0x1.9525a8p-3f, 0x1.e8ad7p-32f, 0.0f},
    {0x1.9ec8eap-1f, 0x1.af3c92p-3f, -0x1.dfdp-29f, 0.0f},
    {0x1.99999ap-1f, 0x1.c8ff7ap-3f, 0x1.e6a68ap-29f, 0.0f},
    {0x1.948b1p-1f, 0x1.e27076p-3f, -0x1.d50d1ap-31f, 0.0f},
    {0x1.8f9c18p-1f, 0x1.fb918cp-3f, -0x1.50e0dep-30f, 0.0f},
    {0x1.8acb9p-1f, 0x1.0a325p-2f, 0x1.4e721ep-27f, 0.0f},
    {0x1.861862p-1f, 0x1.1675cap-2f, -0x1.0a8b3ep-27f, 0.0f},
    {0x1.818182p-1f, 0x1.22941ep-2f, 0x1.b3de5ap-28f, 0.0f},
    {0x1.7d05f4p-1f, 0x1.2e8e2cp-2f, -0x1.1ee2dp-30f, 0.0f},
    {0x1.78a4c8p-1f, 0x1.3a64c6p-2f, -0x1.a5ae86p-28f, 0.0f},
    {0x1.745d18p-1f, 0x1.4618bap-2f, 0x1.0e2f66p-29f, 0.0f},
    {0x1.702e06p-1f, 0x1.51aad8p-2f, -0x1.e903eap-29f, 0.0f},
    {0x1.6c16c2p-1f, 0x1.5d1bdap-2f, 0x1.560274p-28f, 0.0f},
    {0x1.681682p-1f, 0x1.686c8p-2f, 0x1.cd8a5ap-29f, 0.0f},
    {0x1.642c86p-1f, 0x1.739d7ep-2f, 0x1.5de804p-29f, 0.0f},
    {0x1.605816p-1f, 0x1.7eaf84p-2f, -0x1.bea81ep-29f, 0.0f},
    {0x1.5c9882p-1f, 0x1.89a33ap-2f, 0x1.18284cp-27f, 0.0f},
    {0x1.58ed24p-1f, 0x1.94793ep-2f, 0x1.c422e2p-27f, 0.0f},
    {0x1.555556p-1f, 0x1.9f323cp-2f, 0x1.97f30ap-27f, 0.0f},
    {0x1.51d07ep-1f, 0x1.a9ceccp-2f, -0x1.197decp-28f, 0.0f},
    {0x1.4e5e0ap-1f, 0x1.b44f7ap-2f, -0x1.c66e14p-27f, 0.0f},
    {0x1.4afd6ap-1f, 0x1.beb4dap-2f, -0x1.58e484p-30f, 0.0f},
    {0x1.47ae14p-1f, 0x1.c8ff7ep-2f, -0x1.959768p-32f, 0.0f},
    {0x1.446f86p-1f, 0x1.d32fe8p-2f, 0x1.e01d7ap-27f, 0.0f},
    {0x1.414142p-1f, 0x1.dd469ep-2f, -0x1.3e3b54p-30f, 0.0f},
    {0x1.3e22ccp-1f, 0x1.e74426p-2f, -0x1.052f0ep-27f, 0.0f},
    {0x1.3b13b2p-1f, 0x1.f128f4p-2f, -0x1.0a1f2p-27f, 0.0f},
    {0x1.381382p-1f, 0x1.faf586p-2f, 0x1.9e3ccap-28f, 0.0f},
    {0x1.3521dp-1f, 0x1.02552ap-1f, -0x1.2d178p-28f, 0.0f},
    {0x1.323e34p-1f, 0x1.0723e6p-1f, 0x1.a39be8p-26f, 0.0f},
    {0x1.2f684cp-1f, 0x1.0be72ep-1f, 0x1.295418p-32f, 0.0f},
    {0x1.2c9fb4p-1f, 0x1.109f3cp-1f, -0x1.5a566cp-26f, 0.0f},
    {0x1.29e412p-1f, 0x1.154c3ep-1f, 0x1.fa6af8p-28f, 0.0f},
    {0x1.27350cp-1f, 0x1.19ee6ap-1f, 0x1.d9f25cp-27f, 0.0f},
    {0x1.24924ap-1f, 0x1.1e85f4p-1f, 0x1.9c1036p-27f, 0.0f},
    {0x1.21fb78p-1f, 0x1.23130ep-1f, -0x1.90503p-27f, 0.0f},
    {0x1.1f7048p-1f, 0x1.2795ep-1f, 0x1.d13624p-26f, 0.0f},
    {0x1.1cf06ap-1f, 0x1.2c0eap-1f, 0x1.7123a6p-27f, 0.0f},
    {0x1.1a7b96p-1f, 0x1.307d74p-1f, -0x1.561de8p-26f, 0.0f},
    {0x1.181182p-1f, 0x1.34e288p-1f, 0x1.8e70fp-28f, 0.0f},
    {0x1.15b1e6p-1f, 0x1.393e0ep-1f, -0x1.b53abcp-26f, 0.0f},
    {0x1.135c82p-1f, 0x1.3d9024p-1f, 0x1.de2aep-26f, 0.0f},
    {0x1.111112p-1f, 0x1.41d8fcp-1f, 0x1.88ce58p-26f, 0.0f},
    {0x1.0ecf56p-1f, 0x1.4618bep-1f, -0x1.d8e84ep-27f, 0.0f},
    {0x1.0c9714p-1f, 0x1.4a4f88p-1f, -0x1.13f04ep-27f, 0.0f},
    {0x1.0a681p-1f, 0x1.4e7d82p-1f, 0x1.6dd6eep-27f, 0.0f},
    {0x1.08421p-1f, 0x1.52a2d4p-1f, -0x1.34874ap-26f, 0.0f},
    {0x1.0624dep-1f, 0x1.56bf9cp-1f, -0x1.e6062ep-28f, 0.0f},
    {0x1.041042p-1f, 0x1.5ad402p-1f, 0x1.a6b3e8p-26f, 0.0f},
    {0x1.020408p-1f, 0x1.5ee02ap-1f, 0x1.6482cep-26f, 0.0f},
    {0x1p-1f, 0x1.62e43p-1f, -0x1.05c61p-29f, 0.0f},
};

__device__ __forceinline__ void ln_table_to_lds(float4 *lds) {
    for (int k = threadIdx.x; k < kLnCells; k += blockDim.x) lds[k] = kLnU21[k];
}

// ln_u21 in two steps, so a caller can issue the table read early and do
// independent work while it is in flight: the cell (one LDS read), then the
// arithmetic.
struct LnCell {
    float4 c;    // rc, hi, lo
    unsigned b;  // bits of float(n)
};
__device__ __forceinline__ LnCell ln_u21_cell(unsigned n, const float4 *__restrict__ tab) {
    LnCell k;
    k.b = __float_as_uint((float)n);
    k.c = tab[(((k.b >> 16) & 0x7fu) + 1u) >> 1];
    return k;
}
__device__ __forceinline__ float ln_u21_finish(const LnCell &k) {
    const float m = __uint_as_float((k.b & 0x7fffffu) | 0x3f800000u);
    const float ef = (float)((int)(k.b >> 23) - 148);  // E - 21
    const float t = fmaf(m, k.c.x, -1.0f);
    const float t2 = t * t;
    float p = fmaf(t, -0.25f, 0x1.555556p-2f);
    p = fmaf(p, t, -0.5f);
    const float lp = fmaf(t2, p, t);
    const float hi = fmaf(ef, 0x1.62e430p-1f, k.c.y);
    const float lo = fmaf(ef, -0x1.05c610p-29f, k.c.z);
    return hi + (lo + lp);
}
__device__ __forceinline__ float ln_u21(unsigned n, const float4 *__restrict__ tab) {
    return ln_u21_finish(ln_u21_cell(n, tab));
}

// Gaussian angle step of the batched fast path from ONE 64-bit draw x:
//   bits 63..43 -> k1, u1 = (k1 + 1) 2^-21 in (0, 1]
//   bits 42..22 -> sign bit b and 20-bit a: phi = pi (a 2^-20 - 1/2) in [-pi/2, pi/2)
//   bits 21..1  -> k3, acceptance uniform v = k3 2^-21 in [0, 1)
// z = (-1)^b sqrt(-2 log u1) sin(phi): Box-Muller with the sine's sign taken
// from an independent bit, so the proposal is exactly symmetric.  log u1 is
// ln_u21(k1 + 1) with the cell table `tab` (in LDS).
struct GaussDraw {
    float z, v;
};
__device__ __forceinline__ GaussDraw gauss_draw(uint64_t x, const float4 *__restrict__ tab) {
    const unsigned hi = (unsigned)(x >> 32), lo = (unsigned)x;
    const unsigned k1 = hi >> 11;
    const unsigned k2 = ((hi & 0x7ffu) << 10) | (lo >> 22);
    const unsigned k3 = (lo >> 1) & 0x1fffffu;
    const LnCell cell = ln_u21_cell(k1 + 1u, tab);
    // the sine and v while the cell read is in flight
    const float phi = fmaf((float)(k2 & 0xfffffu), 0x1.921fb6p-19f, -0x1.921fb6p+0f);  // pi 2^-20, -pi/2
    const float sn = halfpi_sinf(phi);
    GaussDraw d;
    d.v = (float)k3 * 0x1p-21f;
    __builtin_amdgcn_sched_barrier(0);
    const float z = newton_sqrtf(-2.0f * ln_u21_finish(cell)) * sn;
    d.z = (k2 >> 20) ? -z : z;
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
