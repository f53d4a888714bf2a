// Shared definitions for the gfx950 Metropolis + parallel-tempering library.
//
// Numerical policy (identical in every kernel):
//  * float expressions are evaluated in source order without contraction
//    (the library is compiled with -ffp-contract=off);
//  * a float32 transcendental (cos, sin, exp, log) is this project's float32
//    function of mc_fastmath.h (Cody-Waite reduction and polynomials with
//    explicit fused multiply-adds, at most 1.5 ulp on the arguments the
//    updates pass), which the CPU oracle evaluates identically; the Ising
//    acceptance thresholds exp(-beta dE) are the float64 exponential rounded
//    once to float32 (exact_expf); sqrt and division are IEEE correctly
//    rounded; float64 paths call the float64 functions;
//  * per-site RNG streams are numba's xoroshiro128+ records {s0, s1}.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/mcgpu.h"
#include "mc_fastmath.h"

namespace mc {

// ------------------------------------------------------------ errors
void set_error(int code, const char *fmt, ...);
int fail_hip(hipError_t err, const char *what);

#define MC_TRY_HIP(expr)                                   \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return ::mc::fail_hip(e_, #expr); \
    } while (0)

#define MC_CHECK_LAUNCH() MC_TRY_HIP(hipGetLastError())

#define MC_REQUIRE(cond, ...)                                   \
    do {                                                        \
        if (!(cond)) {                                          \
            ::mc::set_error(MC_ERR_VALUE, __VA_ARGS__);         \
            return MC_ERR_VALUE;                                \
        }                                                       \
    } while (0)

// Library-owned device scratch for reductions, one region per (device,
// stream); grows on demand, never shrinks.  Returns MC_OK and sets *ptr, or
// an error code.
int reduction_scratch(void *stream, size_t bytes, void **ptr);

// Caller-provided partials for the reductions of this thread, in place of
// reduction_scratch (set by the entry points that allocate nothing, for the
// duration of one call).  ptr == nullptr: none.
struct PartialsOverride {
    double *ptr;
    size_t capacity;  // doubles
};
extern thread_local PartialsOverride g_partials_override;

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// ------------------------------------------------------------- RNG
// numba.cuda.random xoroshiro128+ (Blackman & Vigna); record = {s0, s1}.
// Left rotation by k (32 < k < 64) as a right rotation by 64 - k < 32: each
// half is a funnel shift of both halves (v_alignbit), two instructions where
// the shift-and-or form takes three.
__device__ __forceinline__ uint64_t rotl64(uint64_t x, unsigned k) {
    const unsigned lo = (unsigned)x, hi = (unsigned)(x >> 32), r = 64u - k;
    const unsigned nlo = __builtin_amdgcn_alignbit(hi, lo, r), nhi = __builtin_amdgcn_alignbit(lo, hi, r);
    return ((uint64_t)nhi << 32) | nlo;
}

struct Rng {
    uint64_t s0, s1;
    __device__ __forceinline__ uint64_t next() {
        const uint64_t a = s0;
        uint64_t b = s1;
        const uint64_t result = a + b;
        b ^= a;
        s0 = rotl64(a, 55) ^ b ^ (b << 14);
        s1 = rotl64(b, 36);
        return result;
    }
    // (x >> 11) * 2^-53 exactly, as float64
    __device__ __forceinline__ double uniform_f64() {
        return (double)(next() >> 11) * 0x1.0p-53;
    }
    // float64 uniform rounded to nearest float32 (may be 1.0f)
    __device__ __forceinline__ float uniform_f32() { return (float)uniform_f64(); }
    template <class R> __device__ __forceinline__ R uniform();
};
template <> __device__ __forceinline__ float Rng::uniform<float>() { return uniform_f32(); }
template <> __device__ __forceinline__ double Rng::uniform<double>() { return uniform_f64(); }

__device__ __forceinline__ Rng load_rng(const uint64_t *states, size_t k) {
    const ulonglong2 v = reinterpret_cast<const ulonglong2 *>(states)[k];
    Rng g;
    g.s0 = v.x;
    g.s1 = v.y;
    return g;
}
__device__ __forceinline__ void store_rng(uint64_t *states, size_t k, const Rng &g) {
    ulonglong2 v;
    v.x = g.s0;
    v.y = g.s1;
    reinterpret_cast<ulonglong2 *>(states)[k] = v;
}

// ---------------------------------------------------- transcendentals
__device__ __forceinline__ float rcos(float x) {
    float s, c;
    fast_sincosf(x, s, c);
    return c;
}
__device__ __forceinline__ float rsin(float x) {
    float s, c;
    fast_sincosf(x, s, c);
    return s;
}
__device__ __forceinline__ float rexp(float x) { return fast_expf(x); }
__device__ __forceinline__ float rlog(float x) { return fast_logf(x); }
// the Ising thresholds: float64 exp rounded once to float32
__device__ __forceinline__ float exact_expf(float x) { return (float)::exp((double)x); }
__device__ __forceinline__ float rsqrt_(float x) { return ::sqrtf(x); }
__device__ __forceinline__ double rcos(double x) { return ::cos(x); }
__device__ __forceinline__ double rsin(double x) { return ::sin(x); }
__device__ __forceinline__ double rexp(double x) { return ::exp(x); }
__device__ __forceinline__ double rlog(double x) { return ::log(x); }
__device__ __forceinline__ double rsqrt_(double x) { return ::sqrt(x); }

// The reference's Metropolis test "dE <= 0 or u < exp(-beta dE)" with the
// weight evaluated on every lane and the two tests combined without
// short-circuit control flow (the argument is capped at 0; a NaN stays NaN
// and is rejected as before).  hipcc for gfx950 has dropped the accepted
// value of a variable assigned under the short-circuit form once the
// exponential is an inlined function with control flow of its own.
template <class R>
__device__ __forceinline__ bool metropolis_accept(R dE, R u, R beta) {
    R x = -beta * dE;
    x = x > R(0) ? R(0) : x;
    return (dE <= R(0)) | (u < rexp(x));
}

template <class R> struct RealConst;
template <> struct RealConst<float> {
    static __device__ __forceinline__ float pi() { return 3.14159274101257324f; }       // float32(pi)
    static __device__ __forceinline__ float two_pi() { return 6.28318548202514648f; }   // float32(2 pi)
};
template <> struct RealConst<double> {
    static __device__ __forceinline__ double pi() { return 3.141592653589793; }
    static __device__ __forceinline__ double two_pi() { return 6.283185307179586; }
};

// ------------------------------------------------------------ complex
template <class R> struct Cpx {
    R re, im;
};
template <class R> __device__ __forceinline__ Cpx<R> cx(R re, R im) { return Cpx<R>{re, im}; }
template <class R> __device__ __forceinline__ Cpx<R> operator+(Cpx<R> a, Cpx<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R> __device__ __forceinline__ Cpx<R> operator-(Cpx<R> a, Cpx<R> b) { return {a.re - b.re, a.im - b.im}; }
template <class R> __device__ __forceinline__ Cpx<R> operator*(Cpx<R> a, Cpx<R> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class R> __device__ __forceinline__ Cpx<R> scale(R s, Cpx<R> a) { return {s * a.re, s * a.im}; }
template <class R> __device__ __forceinline__ Cpx<R> conj(Cpx<R> a) { return {a.re, -a.im}; }
template <class R> __device__ __forceinline__ Cpx<R> neg(Cpx<R> a) { return {-a.re, -a.im}; }
template <class R> __device__ __forceinline__ R norm2(Cpx<R> z) { return z.re * z.re + z.im * z.im; }

// periodic neighbour index without a modulo on the hot path
__device__ __forceinline__ int wrap_up(int i, int L) { return i == 0 ? L - 1 : i - 1; }
__device__ __forceinline__ int wrap_dn(int i, int L) { return i == L - 1 ? 0 : i + 1; }

// Row staging of the row-staged sweep kernels: LDS rows 0..nrows-1 from the
// global rows (first + lr) mod L of replica block b (row_reals reals per
// row).  A thread takes columns tid, tid + nt, ... of every row; the loads of
// kStageRows rows x kStageCols columns are issued (addresses first, an index
// past the end clamped to a valid one) before any of them is stored, so the
// staging costs a few memory round trips instead of one per value.
constexpr int kStageRows = 3, kStageCols = 4;
template <class R>
__device__ __forceinline__ void stage_rows(R *rows, const R *b, int L, int row_reals, int nrows, int first,
                                           int tid, int nt) {
    const int ncols = (row_reals + nt - 1) / nt;
    for (int lr0 = 0; lr0 < nrows; lr0 += kStageRows)
        for (int c0 = 0; c0 < ncols; c0 += kStageCols) {
            const R *src[kStageRows][kStageCols];
#pragma unroll
            for (int a = 0; a < kStageRows; a++) {
                const int row = lr0 + a < nrows ? lr0 + a : nrows - 1;
                const R *base = b + (size_t)((first + row + L) % L) * row_reals;
#pragma unroll
                for (int c = 0; c < kStageCols; c++) {
                    const int x = tid + (c0 + c) * nt;
                    src[a][c] = base + (x < row_reals ? x : row_reals - 1);
                }
            }
            R v[kStageRows][kStageCols];
#pragma unroll
            for (int a = 0; a < kStageRows; a++)
#pragma unroll
                for (int c = 0; c < kStageCols; c++) v[a][c] = *src[a][c];
#pragma unroll
            for (int a = 0; a < kStageRows; a++)
#pragma unroll
                for (int c = 0; c < kStageCols; c++) {
                    const int x = tid + (c0 + c) * nt;
                    if (lr0 + a < nrows && x < row_reals) rows[(lr0 + a) * row_reals + x] = v[a][c];
                }
        }
}

// The own rows of a row kernel back: LDS rows 1, 3, ..., 2 n - 1 to global
// rows first + 1, first + 3, ... (no wrap: they lie inside the block).
template <class R>
__device__ __forceinline__ void write_rows(R *b, const R *rows, int row_reals, int n, int first, int tid, int nt) {
    const int ncols = (row_reals + nt - 1) / nt;
    for (int kk = 0; kk < n; kk++) {
        R *dst = b + (size_t)(first + 1 + 2 * kk) * row_reals;
        const R *src = rows + (1 + 2 * kk) * row_reals;
        for (int c0 = 0; c0 < ncols; c0 += kStageCols) {
            R v[kStageCols];
#pragma unroll
            for (int c = 0; c < kStageCols; c++) {
                const int x = tid + (c0 + c) * nt;
                v[c] = src[x < row_reals ? x : row_reals - 1];
            }
#pragma unroll
            for (int c = 0; c < kStageCols; c++) {
                const int x = tid + (c0 + c) * nt;
                if (x < row_reals) dst[x] = v[c];
            }
        }
    }
}

// Sum of a wavefront's x in lane 0 by the tree of a shuffle-down reduction
// (lane i adds lane i + 32, then i + 16, ..., then i + 1), with the upper
// lanes brought down by gfx950's permlane swaps (32, 16) and DPP row shifts
// (8, 4, 2, 1) instead of six dependent ds_bpermute round trips.  The same
// additions in the same order: lane 0's sum is bit-identical to the shuffle
// tree's (checked on MI355X over 65,536 wavefronts of values spanning 60
// binades).  Only lane 0's result is meaningful.
template <int N>
__device__ __forceinline__ double dpp_row_shl_f64(double x) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, 0x100 | N, 0xf, 0xf, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), 0x100 | N, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <int HALF>
__device__ __forceinline__ double lanes_above_f64(double x) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
    if (HALF == 32) {
        lo = __builtin_amdgcn_permlane32_swap(lo, lo, false, false)[1];
        hi = __builtin_amdgcn_permlane32_swap(hi, hi, false, false)[1];
    } else {
        lo = __builtin_amdgcn_permlane16_swap(lo, lo, false, false)[1];
        hi = __builtin_amdgcn_permlane16_swap(hi, hi, false, false)[1];
    }
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ double wave_sum(double x) {
    x += lanes_above_f64<32>(x);
    x += lanes_above_f64<16>(x);
    x += dpp_row_shl_f64<8>(x);
    x += dpp_row_shl_f64<4>(x);
    x += dpp_row_shl_f64<2>(x);
    x += dpp_row_shl_f64<1>(x);
    return x;
}

}  // namespace mc
