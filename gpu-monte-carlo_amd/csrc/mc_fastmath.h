// float32 sin/cos, exp and log for the batched fast path, bit-reproducible on
// the host compiler: Cody-Waite reduction (pi/2 in three parts, ln 2 in two)
// with fused multiply-adds, then Taylor polynomials of sin, cos (to x^9,
// x^10) and exp (to x^7) in Horner form; log splits x = 2^e m with
// m in [sqrt(1/2), sqrt(2)) and sums the atanh series of s = (m-1)/(m+1) to
// s^11.  Errors: sin/cos <= 1.5 ulp on [-2 pi, 2 pi], exp <= 1 ulp for
// normal results, log <= 1.5 ulp for normal x.  The Gaussian step of the
// batched XY update samples the normal quantile from a table of cubic pieces.
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
    const int q = (int)k;  // the quadrant is q mod 4
    const float s1 = (q & 1) ? c : s;
    const float c1 = (q & 1) ? s : c;
    // negation flips the sign bit: adding 2^31 does it (the carry leaves the
    // word), one shift-and-add instead of a compare and a select
    sn = __uint_as_float(__float_as_uint(s1) + ((unsigned)(q & 2) << 30));
    cs = __uint_as_float(__float_as_uint(c1) + ((unsigned)((q + 1) & 2) << 30));
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

// accept_weight(x) 2^21, exactly: the scale 2^k of the result built as
// 2^(k+21) (accept_weight's value is at least exp(-87), a normal float, so
// the product by 2^k is exact either way).  v < accept_weight(x) is then
// k3 < accept_weight_k21(x) for v = k3 2^-21.
__device__ __forceinline__ float accept_weight_k21(float x) {
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
    return p * __int_as_float(((int)k + 127 + 21) << 23);
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

// The standard normal quantile for the Gaussian step, as a table of cubic
// pieces.  Input: a 21-bit integer k and the point w = (2k + 1) 2^-23 in
// (0, 1/2); output Q(w) = -Phi^-1(w) > 0 (Phi the normal CDF).  float(2k + 1)
// is exact; with its bits b, cell = 8 E + (top three mantissa bits) =
// (b >> 20) - 1016, where 2^E <= 2k + 1 < 2^(E+1) (E = 0..21, 176 cells,
// eight per octave of w), and t in [0, 1) is the remaining 20 mantissa bits.
// Each cell holds c0..c3 of Q = ((c3 t + c2) t + c1) t + c0, fitted to Q in
// float64 at 64 Chebyshev points of the cell and rounded to float32 (cells
// that no 2k + 1 reaches are fitted the same way).  Error over all 2^21
// inputs: <= 2.01 ulp of max(Q, 1) (|error| < 4.7e-7; max Q = 5.17).  The
// kernels copy the table to LDS (normal_table_to_lds) and index it there.
constexpr int kNormalCells = 176;
static __constant__ float4 kNormalQ[kNormalCells] = {
    {0x1.4aa936p+2f, -0x1.7eed2ep-6f, 0x1.6f09f4p-10f, -0x1.969384p-14f},
    {0x1.493fa4p+2f, -0x1.55c132p-6f, 0x1.238bcep-10f, -0x1.241342p-14f},
    {0x1.47faf8p+2f, -0x1.34b276p-6f, 0x1.da7048p-11f, -0x1.b1d31p-15f},
    {0x1.46d44p+2f, -0x1.198fb6p-6f, 0x1.89a3d4p-11f, -0x1.4b0df2p-15f},
    {0x1.45c658p+2f, -0x1.02e17cp-6f, 0x1.4beb5cp-11f, -0x1.02684cp-15f},
    {0x1.44cd54p+2f, -0x1.df4622p-7f, 0x1.1bb44ap-11f, -0x1.9b30d8p-16f},
    {0x1.43e628p+2f, -0x1.be3372p-7f, 0x1.eaa308p-12f, -0x1.4c8e36p-16f},
    {0x1.430e66p+2f, -0x1.a17858p-7f, 0x1.ac83acp-12f, -0x1.10cd08p-16f},
    {0x1.424418p+2f, -0x1.883e4ap-6f, 0x1.77579ep-10f, -0x1.9f823p-14f},
    {0x1.40d1aep+2f, -0x1.5e2314p-6f, 0x1.2a30bep-10f, -0x1.2a893ap-14f},
    {0x1.3f8504p+2f, -0x1.3c52e8p-6f, 0x1.e552b4p-11f, -0x1.bb7afep-15f},
    {0x1.3e56fep+2f, -0x1.20905ap-6f, 0x1.92b9c4p-11f, -0x1.5276fcp-15f},
    {0x1.3d425ap+2f, -0x1.095bc4p-6f, 0x1.539f8cp-11f, -0x1.0838cap-15f},
    {0x1.3c4318p+2f, -0x1.eb5578p-7f, 0x1.2252b2p-11f, -0x1.a47cep-16f},
    {0x1.3b5616p+2f, -0x1.c97cd2p-7f, 0x1.f6236ep-12f, -0x1.541bc6p-16f},
    {0x1.3a78dcp+2f, -0x1.ac14eep-7f, 0x1.b69ad2p-12f, -0x1.1705e2p-16f},
    {0x1.39a966p+2f, -0x1.924286p-6f, 0x1.80387ap-10f, -0x1.a90a6ap-14f},
    {0x1.382d7ep+2f, -0x1.672778p-6f, 0x1.314c5ap-10f, -0x1.316f08p-14f},
    {0x1.36d83ap+2f, -0x1.448852p-6f, 0x1.f0f8dap-11f, -0x1.c5cb1cp-15f},
    {0x1.35a256p+2f, -0x1.281ab2p-6f, 0x1.9c7414p-11f, -0x1.5a61d8p-15f},
    {0x1.348672p+2f, -0x1.10563cp-6f, 0x1.5bdfe6p-11f, -0x1.0e6facp-15f},
    {0x1.338074p+2f, -0x1.f854dap-7f, 0x1.296a2ap-11f, -0x1.ae6d6ep-16f},
    {0x1.328d28p+2f, -0x1.d5a81p-7f, 0x1.013b9ap-11f, -0x1.5c2f9ep-16f},
    {0x1.31aa08p+2f, -0x1.b786f8p-7f, 0x1.c16c44p-12f, -0x1.1daddp-16f},
    {0x1.30d502p+2f, -0x1.9d113p-6f, 0x1.89bep-10f, -0x1.b33ddp-14f},
    {0x1.2f4edap+2f, -0x1.70e3bp-6f, 0x1.38ecd2p-10f, -0x1.38d19ap-14f},
    {0x1.2df04cp+2f, -0x1.4d6666p-6f, 0x1.fd7a44p-11f, -0x1.d0d6fcp-15f},
    {0x1.2cb1e8p+2f, -0x1.304118p-6f, 0x1.a6e6aep-11f, -0x1.62ddb8p-15f},
    {0x1.2b8e2ep+2f, -0x1.17e216p-6f, 0x1.64bd8p-11f, -0x1.151902p-15f},
    {0x1.2a80e6p+2f, -0x1.033256p-6f, 0x1.310982p-11f, -0x1.b915f8p-16f},
    {0x1.2986cep+2f, -0x1.e2d3d2p-7f, 0x1.07dc3p-11f, -0x1.64d9bp-16f},
    {0x1.289d4ap+2f, -0x1.c3eb8ep-7f, 0x1.cd0f08p-12f, -0x1.24d218p-16f},
    {0x1.27c24p+2f, -0x1.a8c608p-6f, 0x1.93fca4p-10f, -0x1.be30f4p-14f},
    {0x1.2630fcp+2f, -0x1.7b712ap-6f, 0x1.4122c8p-10f, -0x1.40c006p-14f},
    {0x1.24c85cp+2f, -0x1.5704aep-6f, 0x1.05794ep-10f, -0x1.dcb57ep-15f},
    {0x1.2380cp+2f, -0x1.39197ap-6f, 0x1.b228f6p-11f, -0x1.6bfc5cp-15f},
    {0x1.225482p+2f, -0x1.2013e2p-6f, 0x1.6e4c6ap-11f, -0x1.1c42e2p-15f},
    {0x1.213f52p+2f, -0x1.0ad5dcp-6f, 0x1.39423p-11f, -0x1.c48d48p-16f},
    {0x1.203dd6p+2f, -0x1.f124d8p-7f, 0x1.0f02ccp-11f, -0x1.6e2caep-16f},
    {0x1.1f4d6p+2f, -0x1.d165a4p-7f, 0x1.d99e48p-12f, -0x1.2c824p-16f},
    {0x1.1e6bc8p+2f, -0x1.b58264p-6f, 0x1.9f0c8ap-10f, -0x1.c9fbfp-14f},
    {0x1.1cce6cp+2f, -0x1.86ee82p-6f, 0x1.4a01e2p-10f, -0x1.494c0ap-14f},
    {0x1.1b5ad4p+2f, -0x1.617f8p-6f, 0x1.0cc14cp-10f, -0x1.e9817ep-15f},
    {0x1.1a092cp+2f, -0x1.42be5p-6f, 0x1.be569cp-11f, -0x1.75d2aep-15f},
    {0x1.18d3a6p+2f, -0x1.29047ep-6f, 0x1.78a472p-11f, -0x1.23fde6p-15f},
    {0x1.17b5d4p+2f, -0x1.132c76p-6f, 0x1.4228ep-11f, -0x1.d0ee34p-16f},
    {0x1.16ac46p+2f, -0x1.0063d2p-6f, 0x1.16c19ap-11f, -0x1.783ea2p-16f},
    {0x1.15b43ap+2f, -0x1.e01f92p-7f, 0x1.e73a4p-12f, -0x1.34d0a4p-16f},
    {0x1.14cb78p+2f, -0x1.c36eb6p-6f, 0x1.ab0a5ap-10f, -0x1.d6bb42p-14f},
    {0x1.1320e4p+2f, -0x1.9380e6p-6f, 0x1.53a18p-10f, -0x1.528a96p-14f},
    {0x1.11a14ap+2f, -0x1.6cf962p-6f, 0x1.14a892p-10f, -0x1.f75ad2p-15f},
    {0x1.1044ap+2f, -0x1.4d4fdcp-6f, 0x1.cb90a4p-11f, -0x1.807972p-15f},
    {0x1.0f04ecp+2f, -0x1.32d23ep-6f, 0x1.83e1fap-11f, -0x1.2c5dc2p-15f},
    {0x1.0ddda4p+2f, -0x1.1c52d2p-6f, 0x1.4bd64cp-11f, -0x1.de58ap-16f},
    {0x1.0ccb38p+2f, -0x1.08f952p-6f, 0x1.1f2e5ap-11f, -0x1.8329bep-16f},
    {0x1.0bcad8p+2f, -0x1.f04d4p-7f, 0x1.f60982p-12f, -0x1.3dd30ep-16f},
    {0x1.0ada38p+2f, -0x1.d2bc96p-6f, 0x1.b8186ap-10f, -0x1.e490cep-14f},
    {0x1.09211ap+2f, -0x1.a1560cp-6f, 0x1.5e1daap-10f, -0x1.5c94a8p-14f},
    {0x1.079448p+2f, -0x1.799ccep-6f, 0x1.1d4682p-10f, -0x1.0333b6p-14f},
    {0x1.062b7ep+2f, -0x1.58f5e6p-6f, 0x1.d9fecp-11f, -0x1.8c0e3ep-15f},
    {0x1.04e092p+2f, -0x1.3da29cp-6f, 0x1.90272ep-11f, -0x1.3579fep-15f},
    {0x1.03aed6p+2f, -0x1.266c6cp-6f, 0x1.56683cp-11f, -0x1.ecf2aap-16f},
    {0x1.0292ap+2f, -0x1.1274acp-6f, 0x1.28634p-11f, -0x1.8f0d6p-16f},
    {0x1.01890cp+2f, -0x1.011784p-6f, 0x1.031d52p-11f, -0x1.47a3a4p-16f},
    {0x1.008fbap+2f, -0x1.e3a9a2p-6f, 0x1.c6603ap-10f, -0x1.f3a546p-14f},
    {0x1.fd8d06p+1f, -0x1.b0a6d6p-6f, 0x1.69984ep-10f, -0x1.678842p-14f},
    {0x1.fa561ep+1f, -0x1.879ec6p-6f, 0x1.26b774p-10f, -0x1.0b6a7ep-14f},
    {0x1.f769ap+1f, -0x1.65e226p-6f, 0x1.e9d118p-11f, -0x1.98b4b2p-15f},
    {0x1.f4baep+1f, -0x1.49a486p-6f, 0x1.9d9d92p-11f, -0x1.3f6fp-15f},
    {0x1.f24032p+1f, -0x1.31a5bep-6f, 0x1.6202e4p-11f, -0x1.fcea6p-16f},
    {0x1.eff208p+1f, -0x1.1d0036p-6f, 0x1.328038p-11f, -0x1.9c0f6cp-16f},
    {0x1.edca62p+1f, -0x1.0b0aeep-6f, 0x1.0c0332p-11f, -0x1.5262p-16f},
    {0x1.ebc462p+1f, -0x1.f6836cp-6f, 0x1.d61462p-10f, -0x1.0214f4p-13f},
    {0x1.e80e16p+1f, -0x1.c1bb16p-6f, 0x1.763aep-10f, -0x1.7389cap-14f},
    {0x1.e4b68p+1f, -0x1.974258p-6f, 0x1.311e16p-10f, -0x1.146d7ap-14f},
    {0x1.e1abf6p+1f, -0x1.7453aap-6f, 0x1.fb42aep-11f, -0x1.a6982p-15f},
    {0x1.dee15cp+1f, -0x1.571398p-6f, 0x1.ac7814p-11f, -0x1.4a5f5p-15f},
    {0x1.dc4cb2p+1f, -0x1.3e376ep-6f, 0x1.6ed2c2p-11f, -0x1.073beap-15f},
    {0x1.d9e62ap+1f, -0x1.28d1f6p-6f, 0x1.3dac8p-11f, -0x1.aa5e1ap-16f},
    {0x1.d7a78ap+1f, -0x1.163492p-6f, 0x1.15d94cp-11f, -0x1.5e34b6p-16f},
    {0x1.d58bdp+1f, -0x1.05d69ap-5f, 0x1.e7733cp-10f, -0x1.0b2d7p-13f},
    {0x1.d1ad38p+1f, -0x1.d4eef6p-6f, 0x1.843886p-10f, -0x1.80c5cap-14f},
    {0x1.ce30dep+1f, -0x1.a8ddcep-6f, 0x1.3ca53ep-10f, -0x1.1e5ea6p-14f},
    {0x1.cb047cp+1f, -0x1.849bc6p-6f, 0x1.074e48p-10f, -0x1.b5edb2p-15f},
    {0x1.c81a78p+1f, -0x1.663cep-6f, 0x1.bcf5d4p-11f, -0x1.567566p-15f},
    {0x1.c56876p+1f, -0x1.4c6ad4p-6f, 0x1.7d0f04p-11f, -0x1.10f002p-15f},
    {0x1.c2e662p+1f, -0x1.36300ap-6f, 0x1.4a18d8p-11f, -0x1.ba324ep-16f},
    {0x1.c08dc6p+1f, -0x1.22d7aap-6f, 0x1.20cb1p-11f, -0x1.6b4b54p-16f},
    {0x1.be596cp+1f, -0x1.11d426p-5f, 0x1.faca62p-10f, -0x1.154142p-13f},
    {0x1.ba4d2p+1f, -0x1.eabaeep-6f, 0x1.93d10ap-10f, -0x1.8f7356p-14f},
    {0x1.b6a706p+1f, -0x1.bce248p-6f, 0x1.49827ap-10f, -0x1.296818p-14f},
    {0x1.b3542p+1f, -0x1.972564p-6f, 0x1.121d0cp-10f, -0x1.c6f756p-15f},
    {0x1.b04652p+1f, -0x1.7785fap-6f, 0x1.cf65e2p-11f, -0x1.63e5f6p-15f},
    {0x1.ad72d8p+1f, -0x1.5ca0ecp-6f, 0x1.8cfce4p-11f, -0x1.1bbc5ep-15f},
    {0x1.aad14ap+1f, -0x1.457758p-6f, 0x1.580274p-11f, -0x1.cbd2e6p-16f},
    {0x1.a85af6p+1f, -0x1.314d7cp-6f, 0x1.2d0f14p-11f, -0x1.79e11cp-16f},
    {0x1.a60a6ep+1f, -0x1.1f906p-5f, 0x1.083dcp-9f, -0x1.207ec2p-13f},
    {0x1.a1c9bap+1f, -0x1.01dff6p-5f, 0x1.a554d8p-10f, -0x1.9fd74cp-14f},
    {0x1.9df3a6p+1f, -0x1.d3e774p-6f, 0x1.57f962p-10f, -0x1.35be84p-14f},
    {0x1.9a746ap+1f, -0x1.ac805cp-6f, 0x1.1e476cp-10f, -0x1.da07bep-15f},
    {0x1.973d58p+1f, -0x1.8b77fep-6f, 0x1.e42c5ep-11f, -0x1.72f338p-15f},
    {0x1.944338p+1f, -0x1.6f5ceep-6f, 0x1.9ef428p-11f, -0x1.27d6eap-15f},
    {0x1.917d46p+1f, -0x1.5725e8p-6f, 0x1.67b704p-11f, -0x1.df990ap-16f},
    {0x1.8ee486p+1f, -0x1.420f7ep-6f, 0x1.3aeab4p-11f, -0x1.8a40bcp-16f},
    {0x1.8c735p+1f, -0x1.2f809ep-5f, 0x1.148164p-9f, -0x1.2d2074p-13f},
    {0x1.87f5b8p+1f, -0x1.106d36p-5f, 0x1.b92a62p-10f, -0x1.b248b8p-14f},
    {0x1.83e7c6p+1f, -0x1.eebe04p-6f, 0x1.68602ep-10f, -0x1.43a4aap-14f},
    {0x1.8034cep+1f, -0x1.c5735ep-6f, 0x1.2c172ep-10f, -0x1.ef87c2p-15f},
    {0x1.7ccd7ap+1f, -0x1.a2d1p-6f, 0x1.fbc946p-11f, -0x1.83f138p-15f},
    {0x1.79a612p+1f, -0x1.855568p-6f, 0x1.b36518p-11f, -0x1.358436p-15f},
    {0x1.76b568p+1f, -0x1.6bebbp-6f, 0x1.799a04p-11f, -0x1.f5f638p-16f},
    {0x1.73f42ep+1f, -0x1.55c7bcp-6f, 0x1.4ab6ccp-11f, -0x1.9cc956p-16f},
    {0x1.715c7cp+1f, -0x1.4249a6p-5f, 0x1.227fbap-9f, -0x1.3b70fep-13f},
    {0x1.6c9708p+1f, -0x1.21a0e8p-5f, 0x1.cfd51ep-10f, -0x1.c736e4p-14f},
    {0x1.6846fp+1f, -0x1.07471cp-5f, 0x1.7b25ap-10f, -0x1.536ff6p-14f},
    {0x1.645692p+1f, -0x1.e319d2p-6f, 0x1.3beb22p-10f, -0x1.03fee8p-14f},
    {0x1.60b5d4p+1f, -0x1.bea154p-6f, 0x1.0b7092p-10f, -0x1.974bdep-15f},
    {0x1.5d5868p+1f, -0x1.9f910ap-6f, 0x1.cae01ap-11f, -0x1.451c5ep-15f},
    {0x1.5a34aep+1f, -0x1.84c6dap-6f, 0x1.8e2b62p-11f, -0x1.07be3ap-15f},
    {0x1.5742fcp+1f, -0x1.6d6cd2p-6f, 0x1.5ce5a6p-11f, -0x1.b1f572p-16f},
    {0x1.547d16p+1f, -0x1.58db46p-5f, 0x1.329f2ap-9f, -0x1.4bd0a4p-13f},
    {0x1.4f6122p+1f, -0x1.365f2ap-5f, 0x1.e9fe08p-10f, -0x1.df318p-14f},
    {0x1.4ac126p+1f, -0x1.1a866ap-5f, 0x1.90d7e6p-10f, -0x1.658ee6p-14f},
    {0x1.46865ep+1f, -0x1.038cp-5f, 0x1.4e3ce4p-10f, -0x1.120bf8p-14f},
    {0x1.429fdp+1f, -0x1.e07f12p-6f, 0x1.1b2358p-10f, -0x1.ad8f06p-15f},
    {0x1.3f008ap+1f, -0x1.bf998ep-6f, 0x1.e61dc6p-11f, -0x1.5711bp-15f},
    {0x1.3b9e62p+1f, -0x1.a3363ep-6f, 0x1.a60e5cp-11f, -0x1.1674a8p-15f},
    {0x1.38714p+1f, -0x1.8a73fp-6f, 0x1.7208cap-11f, -0x1.ca643ap-16f},
    {0x1.357292p+1f, -0x1.74a20cp-5f, 0x1.4561bep-9f, -0x1.5ebc9p-13f},
    {0x1.2febe8p+1f, -0x1.5005cp-5f, 0x1.043d54p-9f, -0x1.faf3ap-14f},
    {0x1.2ae8eap+1f, -0x1.326e6ap-5f, 0x1.aa29bap-10f, -0x1.7a916cp-14f},
    {0x1.26518p+1f, -0x1.19fe1cp-5f, 0x1.63a44ep-10f, -0x1.225c86p-14f},
    {0x1.2213b8p+1f, -0x1.05737p-5f, 0x1.2d7fa6p-10f, -0x1.c7712p-15f},
    {0x1.1e21d2p+1f, -0x1.e7dc58p-6f, 0x1.0300b6p-10f, -0x1.6bf95ap-15f},
    {0x1.1a710ep+1f, -0x1.c999ecp-6f, 0x1.c20a6p-11f, -0x1.279a2ap-15f},
    {0x1.16f8d4p+1f, -0x1.af3168p-6f, 0x1.8ad082p-11f, -0x1.e6e5a4p-16f},
    {0x1.13b22ap+1f, -0x1.97e7dep-5f, 0x1.5b639p-9f, -0x1.74d86p-13f},
    {0x1.0da39p+1f, -0x1.70cceep-5f, 0x1.16232ap-9f, -0x1.0db8f8p-13f},
    {0x1.0821aep+1f, -0x1.51291cp-5f, 0x1.c7e8d6p-10f, -0x1.9333eep-14f},
    {0x1.0312ep+1f, -0x1.370178p-5f, 0x1.7ccbfcp-10f, -0x1.358528p-14f},
    {0x1.fcc812p+0f, -0x1.2100dp-5f, 0x1.431458p-10f, -0x1.e5e15p-15f},
    {0x1.f40d06p+0f, -0x1.0e38e4p-5f, 0x1.15be0ep-10f, -0x1.849716p-15f},
    {0x1.ebdda4p+0f, -0x1.fbfc58p-6f, 0x1.e2ebeep-11f, -0x1.3bd14ap-15f},
    {0x1.e4279ap+0f, -0x1.dfa3d6p-6f, 0x1.a7eaecp-11f, -0x1.04458ep-15f},
    {0x1.dcdbfcp+0f, -0x1.c6a14p-5f, 0x1.75383p-9f, -0x1.8efbbp-13f},
    {0x1.cf5518p+0f, -0x1.9c9946p-5f, 0x1.2b1bf6p-9f, -0x1.20f80ep-13f},
    {0x1.c2fcd4p+0f, -0x1.7a8f84p-5f, 0x1.eaafd2p-10f, -0x1.b070ep-14f},
    {0x1.b79c42p+0f, -0x1.5e66e2p-5f, 0x1.9a1fdp-10f, -0x1.4c4cap-14f},
    {0x1.ad0a62p+0f, -0x1.46b2dcp-5f, 0x1.5c283cp-10f, -0x1.0510fp-14f},
    {0x1.a327c2p+0f, -0x1.3274acp-5f, 0x1.2b6f9ep-10f, -0x1.a1f514p-15f},
    {0x1.99dbb4p+0f, -0x1.20f4b8p-5f, 0x1.046954p-10f, -0x1.53f97ap-15f},
    {0x1.91128p+0f, -0x1.11ab4p-5f, 0x1.c94de2p-11f, -0x1.18689ap-15f},
    {0x1.88bc1ep+0f, -0x1.042d96p-4f, 0x1.92b0a2p-9f, -0x1.ae5932p-13f},
    {0x1.79352ap+0f, -0x1.db0134p-5f, 0x1.42bfacp-9f, -0x1.38259ep-13f},
    {0x1.6af4cp+0f, -0x1.b647dp-5f, 0x1.08aacep-9f, -0x1.d3c9b6p-14f},
    {0x1.5dbf88p+0f, -0x1.97e95cp-5f, 0x1.ba2ec6p-10f, -0x1.67f4b4p-14f},
    {0x1.516928p+0f, -0x1.7e5d8ap-5f, 0x1.770e2cp-10f, -0x1.1b2dcp-14f},
    {0x1.45cf94p+0f, -0x1.6891e4p-5f, 0x1.423498p-10f, -0x1.c5f9f6p-15f},
    {0x1.3ad806p+0f, -0x1.55c072p-5f, 0x1.17d0f8p-10f, -0x1.71c8a6p-15f},
    {0x1.306d12p+0f, -0x1.4556b4p-5f, 0x1.ea8ad2p-11f, -0x1.316dap-15f},
    {0x1.267d4ap+0f, -0x1.36e172p-4f, 0x1.af11e4p-9f, -0x1.d5c90ep-13f},
    {0x1.13d80ep+0f, -0x1.1ea8eep-4f, 0x1.57cb48p-9f, -0x1.55ddaep-13f},
    {0x1.028eb6p+0f, -0x1.0b27aep-4f, 0x1.182d34p-9f, -0x1.011d48p-13f},
    {0x1.e4c092p-1f, -0x1.f645b2p-5f, 0x1.d08a44p-10f, -0x1.8d650cp-14f},
    {0x1.c63812p-1f, -0x1.db8beap-5f, 0x1.866c78p-10f, -0x1.3a35ecp-14f},
    {0x1.a938b8p-1f, -0x1.c4f8ap-5f, 0x1.4bc6c6p-10f, -0x1.fabb1p-15f},
    {0x1.8d8726p-1f, -0x1.b1b564p-5f, 0x1.1c74e6p-10f, -0x1.9face2p-15f},
    {0x1.72f38cp-1f, -0x1.a123aap-5f, 0x1.eb3d8p-11f, -0x1.5a37b6p-15f},
    {0x1.5956b4p-1f, -0x1.92c6fcp-4f, 0x1.a7f71ep-9f, -0x1.0e68dp-12f},
    {0x1.2884p-1f, -0x1.7b6a2ap-4f, 0x1.4373a2p-9f, -0x1.94a31cp-13f},
    {0x1.f481cap-2f, -0x1.698c5p-4f, 0x1.f03d2p-10f, -0x1.3c2d7ep-13f},
    {0x1.9be76ep-2f, -0x1.5be0f6p-4f, 0x1.7a5c46p-10f, -0x1.014e38p-13f},
    {0x1.464964p-2f, -0x1.518d8ep-4f, 0x1.1a569ap-10f, -0x1.b3c836p-14f},
    {0x1.e5ca36p-3f, -0x1.4a001p-4f, 0x1.91ef3p-11f, -0x1.80233p-14f},
    {0x1.422c1ap-3f, -0x1.44d772p-4f, 0x1.0268dp-11f, -0x1.60d38ap-14f},
    {0x1.412d5ep-4f, -0x1.41d5f6p-4f, 0x1.fa318cp-13f, -0x1.520752p-14f},
};

__device__ __forceinline__ void normal_table_to_lds(float4 *lds) {
    for (int k = threadIdx.x; k < kNormalCells; k += blockDim.x) lds[k] = kNormalQ[k];
}

// The table for gauss_draw_k: each cell's cubic in the 20 mantissa bits m of
// w themselves (coefficients of t, t^2, t^3 scaled by 2^-20, 2^-40, 2^-60,
// exact).  Every step of the Horner evaluation is then the original's scaled
// by a power of two, and rounding commutes with such scaling while no value
// comes near the subnormal range, so z is the same bit for bit (checked over
// all 2^21 inputs: the smallest scaled coefficient is 1.5e-23, the smallest
// scaled partial sum 1.5e-16).
__device__ __forceinline__ void normal_table_k_to_lds(float4 *lds) {
    for (int k = threadIdx.x; k < kNormalCells; k += blockDim.x) {
        const float4 c = kNormalQ[k];
        lds[k] = make_float4(c.x, c.y * 0x1p-20f, c.z * 0x1p-40f, c.w * 0x1p-60f);
    }
}

// Gaussian angle step of the batched fast path from ONE 64-bit draw x:
//   bits 63..43 -> k, z = (-1)^s Q((2k + 1) 2^-23) (inverse-CDF sampling)
//   bit 42      -> the sign s, so the proposal is exactly symmetric
//   bits 21..1  -> k3, acceptance uniform v = k3 2^-21 in [0, 1)
// `tab` is the quantile table in LDS.
struct GaussDraw {
    float z, v;
};
__device__ __forceinline__ GaussDraw gauss_draw(uint64_t x, const float4 *__restrict__ tab) {
    const unsigned hi = (unsigned)(x >> 32), lo = (unsigned)x;
    const unsigned b = __float_as_uint((float)(2u * (hi >> 11) + 1u));
    const float4 c = tab[(b >> 20) - 1016u];
    const float t = (float)(b & 0xfffffu) * 0x1p-20f;
    const float z = fmaf(fmaf(fmaf(c.w, t, c.z), t, c.y), t, c.x);
    GaussDraw d;
    d.z = (hi & 0x400u) ? -z : z;
    d.v = (float)((lo >> 1) & 0x1fffffu) * 0x1p-21f;
    return d;
}

// gauss_draw with the table of normal_table_k_to_lds (no scaling of t) and
// the acceptance uniform as the integer k3 of v = k3 2^-21 (exact in float),
// to be compared with accept_weight_k21: the same z and the same decisions,
// two instructions fewer per draw.
__device__ __forceinline__ GaussDraw gauss_draw_k(uint64_t x, const float4 *__restrict__ tab) {
    const unsigned hi = (unsigned)(x >> 32), lo = (unsigned)x;
    const unsigned b = __float_as_uint((float)(2u * (hi >> 11) + 1u));
    const float4 c = tab[(b >> 20) - 1016u];
    const float t = (float)(b & 0xfffffu);
    const float z = fmaf(fmaf(fmaf(c.w, t, c.z), t, c.y), t, c.x);
    GaussDraw d;
    // Every table value is positive (the smallest, at k = 2^21 - 1, is
    // 2.9e-7), so negating is setting the sign bit: bit 10 of hi moved to
    // bit 31 and or-ed in, two instructions instead of a compare and a select.
    d.z = __uint_as_float(__float_as_uint(z) | ((hi << 21) & 0x80000000u));
    d.v = (float)((lo >> 1) & 0x1fffffu);
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
