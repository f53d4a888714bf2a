// float32 sin/cos and exp for the batched fast path, bit-reproducible on the
// host compiler: Cody-Waite reduction (pi/2 in three parts, ln 2 in two)
// with fused multiply-adds, then Taylor polynomials of sin, cos (to x^9,
// x^10) and exp (to x^7) in Horner form.  Errors: sin/cos <= 1.5 ulp on
// [-2 pi, 2 pi], exp <= 1 ulp for normal results.
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

__device__ __forceinline__ float fast_expf(float x) {
    if (!(x > -104.0f)) return x != x ? x : 0.0f;
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
    int e = (int)k;
    if (e < -125) {
        p *= 0x1p-64f;
        e += 64;
    }
    return p * __int_as_float((e + 127) << 23);
}

}  // namespace mc
