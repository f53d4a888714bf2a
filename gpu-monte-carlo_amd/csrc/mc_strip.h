// Strip machinery shared by the persistent engines (mc_batch.hip keeps every
// strip in LDS for a whole launch; mc_pack.hip streams strips of several
// lattice sizes through LDS every sweep).
//
// A strip is H full rows of one replica lattice (L x L).  Its LDS image holds
//   local row 0      bottom row of the strip above (received)
//   rows 1..H        own rows: state, RNG streams, cos/sin (or spins)
//   row H+1          ghost: the top row of the strip below, with a copy of its
//                    RNG streams, updated in lock-step with its owner
//   row H+2          second row of the strip below (received)
// with every row stored colour-split (even columns, then odd columns), so a
// colour step reads and writes unit-stride.
//
// Energies are accumulated as integers: Ising bond sums exactly, XY bond
// terms (float32, formed as the kernels form them) rounded to multiples of
// 2^-40 and added in int64.  The sums are therefore independent of how sites
// are spread over threads, strips and workgroups, and exchange decisions do
// not depend on the launch shape.  |bond term| < 4 and at most 2^20 sites per
// replica keep every sum below 2^62.
#pragma once

#include "mc_common.h"
#include "mc_fastmath.h"

namespace mc {

enum { MODEL_ISING = 0, MODEL_XY_FAST = 1, MODEL_XY_GAUSS = 2 };
enum { BATCH_ERR_WATCHDOG = 1 };

constexpr unsigned kWatchdog = 1u << 20;  // about a second of polling
constexpr int kChannels = 5;               // XY: e, mx, my, cx, qx ; Ising: e, m

// LDS position of column j in LDS row `row` (colour-split layout)
__device__ __forceinline__ int lds_index(int row, int j, int L) {
    return row * L + (j & 1) * (L >> 1) + (j >> 1);
}

// One group of colour rows of a pass: colour c, colour rows lo..hi (empty if
// hi < lo).  Colour rows of colours 0/1 are local rows 1, 3, ..., H+1
// (ri = 0..H/2, ri = H/2 being the ghost row); of colours 2/3 local rows
// 2, 4, ..., H (ri = 0..H/2-1).
struct Seg {
    int c, lo, hi;
};

// XY bond term x rounded to the nearest multiple of 2^-40 (ties to even), as
// an integer: x * 2^40 is exact in float64, and adding 1.5 * 2^52 rounds it
// to an integer held in the low mantissa bits (|x| < 2^11).
__device__ __forceinline__ long long bond_fixed(float x) {
    const double magic = 0x1.8p52;
    const double t = (double)x * 0x1p40 + magic;
    return __double_as_longlong(t) - __double_as_longlong(magic);
}

// Energy of an integer energy sum: Ising -J * sum s (s_down + s_right);
// XY -J * 2^-40 * sum of rounded (cos bond x + cos bond y).
__host__ __device__ inline double energy_from_fixed(int model, float J, long long e) {
    return model == MODEL_ISING ? -(double)J * (double)e : -(double)J * ((double)e * 0x1p-40);
}

// Cross-workgroup words (halo and energy words) are written and read with
// agent-scope relaxed atomics, which are device-coherent on their own (gfx950
// sets the device-scope cache bits on them), so no L2 write-back or
// invalidate is needed.  Each word is a single-copy atomic 64-bit value whose
// high half names the sweep or exchange it belongs to, so a reader that sees
// the expected tag also sees the payload written with it.
//
// One thread: poll a word until it carries tag `seq`.  False once the
// watchdog expires or another workgroup has recorded an error.
__device__ inline bool poll_word(const uint64_t *p, unsigned seq, uint32_t &payload, const int *error,
                                 unsigned limit = kWatchdog) {
    unsigned n = 0;
    for (;;) {
        const uint64_t w = __hip_atomic_load(const_cast<uint64_t *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((unsigned)(w >> 32) == seq) {
            payload = (uint32_t)w;
            return true;
        }
        __builtin_amdgcn_s_sleep(1);
        if ((++n & 1023u) == 0 &&
            (n > limit || __hip_atomic_load(const_cast<int *>(error), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)))
            return false;
    }
}

__device__ __forceinline__ void publish_word(uint64_t *p, unsigned seq, uint32_t payload) {
    __hip_atomic_store(p, ((uint64_t)seq << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// x of the lane SH below in the same row of 16 lanes (0 for the first SH
// lanes of a row): a DPP row shift of each 32-bit half.
template <int SH>
__device__ __forceinline__ long long dpp_row_shr_i64(long long x) {
    const unsigned long long u = (unsigned long long)x;
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, 0x110 | SH, 0xf, 0xf, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), 0x110 | SH, 0xf, 0xf, true);
    return (long long)(((unsigned long long)hi << 32) | lo);
}

// Sum of x over the 64 lanes of a wavefront, in every lane; every lane must
// be active.  Inclusive sums within rows of 16 lanes by DPP row shifts, then
// the four row totals from lanes 15, 31, 47 and 63.  A shuffle reduction
// (ds_bpermute) takes six dependent LDS round trips instead.
__device__ __forceinline__ long long wave_sum_i64(long long x) {
    x += dpp_row_shr_i64<1>(x);
    x += dpp_row_shr_i64<2>(x);
    x += dpp_row_shr_i64<4>(x);
    x += dpp_row_shr_i64<8>(x);
    const unsigned long long u = (unsigned long long)x;
    long long t = 0;
#pragma unroll
    for (int r = 15; r < 64; r += 16) {
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, r);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), r);
        t += (long long)(((unsigned long long)hi << 32) | lo);
    }
    return t;
}

// Ising acceptance threshold for 2 s sum_nn = v: always flip when dE <= 0,
// else u < exp(-beta dE) in float32 through the float64 exponential.
__device__ __forceinline__ float ising_threshold(int v, float beta, float J) {
    const float dE = (float)v * J;
    if (dE <= 0.0f) return __builtin_inff();
    return exact_expf(-beta * dE);
}

// Block reduction of one int64 and C float64 channels over NT threads; the
// result is valid in thread 0 (waves added in order).
template <int C, int NT>
__device__ void block_sum(long long &e, double (&v)[C], long long *she, double (*sh)[NT / 64]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    e = wave_sum_i64(e);  // every lane (integers: any order)
    if (lane == 0) she[wave] = e;
#pragma unroll
    for (int c = 0; c < C; c++) {
        const double x = wave_sum(v[c]);  // lane 0: the shuffle-down tree's sum
        if (lane == 0) sh[c][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        long long te = she[0];
#pragma unroll
        for (int w = 1; w < NT / 64; w++) te += she[w];
        e = te;
#pragma unroll
        for (int c = 0; c < C; c++) {
            double t = sh[c][0];
#pragma unroll
            for (int w = 1; w < NT / 64; w++) t += sh[c][w];
            v[c] = t;
        }
    }
    __syncthreads();
}

// LDS bytes of one strip's site data: RNG streams (when staged in LDS) and
// angles of rows 1..H+1, cos/sin (or spins) of rows 0..H+2
__host__ __device__ inline size_t strip_bytes(int model, int L, int H, bool rng_in_lds = true) {
    const size_t rng = rng_in_lds ? (size_t)(H + 1) * L * 16 : 0;
    return model == MODEL_ISING ? rng + (size_t)(H + 3) * L
                                : rng + (size_t)(H + 3) * L * 8 + (size_t)(H + 1) * L * 4;
}

// LDS arrays of one strip
struct StripView {
    ulonglong2 *rng;  // rows 1..H+1 (when staged in LDS)
    float2 *cs;       // rows 0..H+2 (XY): (cos, sin) of each site, one 8-byte LDS access
    float *th;        // rows 1..H+1 (XY)
    int8_t *sp;       // rows 0..H+2 (Ising)
};

__device__ __forceinline__ StripView strip_view(unsigned char *base, int L, int H, bool rng_in_lds = true) {
    StripView v;
    v.rng = reinterpret_cast<ulonglong2 *>(base);
    unsigned char *site = base + (rng_in_lds ? (size_t)(H + 1) * L * 16 : 0);
    v.cs = reinterpret_cast<float2 *>(site);
    v.th = reinterpret_cast<float *>(v.cs + (H + 3) * L);
    v.sp = reinterpret_cast<int8_t *>(site);
    return v;
}

// A site's draws for one sweep, taken from its stream before the sweep's
// colour passes (they do not depend on the lattice): Ising a = u; XY a = u1
// (proposal), b = u2 (acceptance); Gaussian XY a = z, b = v (gauss_draw with
// the normal quantile table q_tab in LDS; unused by the other models).
struct SiteDraw {
    float a, b;
};

template <int MODEL>
__device__ __forceinline__ SiteDraw site_draw(Rng &g, const float4 *q_tab) {
    SiteDraw d;
    if (MODEL == MODEL_XY_GAUSS) {
        const GaussDraw gd = gauss_draw(g.next(), q_tab);
        d.a = gd.z;
        d.b = gd.v;
    } else if (MODEL == MODEL_XY_FAST) {
        d.a = g.uniform_f32();
        d.b = g.uniform_f32();
    } else {
        d.a = g.uniform_f32();
        d.b = 0.0f;
    }
    return d;
}

// The update of site_update (below) with the site's draws given and its LDS
// positions precomputed: self (cos/sin or spin arrays), lf and rt (left and
// right neighbours); up and down are self -/+ L, the angle is at self - L.
template <int MODEL>
__device__ __forceinline__ bool site_apply(const StripView &v, const int self, const int lf, const int rt,
                                           const int L, const float beta, const float sigma, const float J,
                                           const float delta, const float *thr, const SiteDraw d) {
    const int up = self - L, dn = self + L, own = self - L;
    if (MODEL == MODEL_ISING) {
        const int sv = v.sp[self];
        const int nb = (int)v.sp[up] + (int)v.sp[dn] + (int)v.sp[lf] + (int)v.sp[rt];
        const bool acc = d.a < thr[(2 * sv * nb + 8) >> 2];
        if (acc) v.sp[self] = (int8_t)(-sv);
        return acc;
    } else {
        const float th = v.th[own];
        float pr;
        if (MODEL == MODEL_XY_GAUSS) {
            pr = wrap_turns(th + sigma * d.a);
        } else {
            const float pi = 3.14159274101257324f, two_pi = 6.28318548202514648f;
            pr = th + delta * (2.0f * d.a - 1.0f);
            if (pr >= pi)
                pr -= two_pi;
            else if (pr < -pi)
                pr += two_pi;
        }
        float sp, cp;
        fast_sincosf(pr, sp, cp);
        const float2 nu = v.cs[up], nd = v.cs[dn], nl = v.cs[lf], nr = v.cs[rt], no = v.cs[self];
        const float Cs = ((nu.x + nd.x) + nl.x) + nr.x;
        const float Ss = ((nu.y + nd.y) + nl.y) + nr.y;
        float t = (cp - no.x) * Cs;
        t = fmaf(sp - no.y, Ss, t);
        const float dE = -J * t;
        const bool acc = d.b < accept_weight(-beta * dE);  // dE <= 0 or b < exp(-beta dE)
        if (acc) {
            v.th[own] = pr;
            v.cs[self] = make_float2(cp, sp);
        }
        return acc;
    }
}

// Update site (colour row ri, column 2 tcol + (c & 1)) of colour c of a
// strip; true if the move was accepted.
//  * MODEL_ISING: the reference's update (models/ising/update.py:16-29) with
//    thresholds thr[k] for 2 s sum_nn = 4k - 8.
//  * MODEL_XY_FAST: the reference's proposal and rule (models/xy/update.py:
//    21-54) with dE from the cached (cos, sin) of the neighbours and the
//    float32 sincos/exp of mc_fastmath.h.
//  * MODEL_XY_GAUSS: proposal theta + sigma z; z and the acceptance uniform
//    from one 64-bit draw (gauss_draw, quantile table q_tab in LDS).
template <int MODEL>
__device__ __forceinline__ bool site_update(const StripView &v, const int c, const int ri, const int tcol,
                                            const int L, const float beta, const float sigma, const float J,
                                            const float delta, const float *thr, const float4 *q_tab) {
    const int half = L >> 1;
    const int p = c & 1;
    const int lr = 1 + 2 * ri + (c >> 1);
    const int base = lr * L;
    const int self = base + p * half + tcol;
    const int up = self - L, dn = self + L;
    const int lf = p ? base + tcol : base + half + (tcol == 0 ? half - 1 : tcol - 1);
    const int rt = p ? base + (tcol == half - 1 ? 0 : tcol + 1) : base + half + tcol;
    const int own = self - L;  // RNG and angle arrays start at local row 1
    const ulonglong2 st = v.rng[own];
    Rng g;
    g.s0 = st.x;
    g.s1 = st.y;
    bool acc;
    if (MODEL == MODEL_ISING) {
        const int sv = v.sp[self];
        const int nb = (int)v.sp[up] + (int)v.sp[dn] + (int)v.sp[lf] + (int)v.sp[rt];
        const float u = g.uniform_f32();
        acc = u < thr[(2 * sv * nb + 8) >> 2];
        if (acc) v.sp[self] = (int8_t)(-sv);
    } else {
        const float one = 1.0f, two = 2.0f;
        const float pi = 3.14159274101257324f, two_pi = 6.28318548202514648f;
        const float th = v.th[own];
        // the site's and its neighbours' (cos, sin) are read before the draw,
        // so their LDS latency runs under the proposal's arithmetic
        const float2 u2 = v.cs[up], d2 = v.cs[dn], l2 = v.cs[lf], r2 = v.cs[rt], o2 = v.cs[self];
        const float cu = u2.x, cd = d2.x, cl = l2.x, cr = r2.x, c0 = o2.x;
        const float su = u2.y, sd = d2.y, sl = l2.y, sr = r2.y, s0 = o2.y;
        __builtin_amdgcn_sched_barrier(0);
        float pr;
        float w;
        if (MODEL == MODEL_XY_GAUSS) {
            const GaussDraw d = gauss_draw(g.next(), q_tab);
            pr = wrap_turns(th + sigma * d.z);
            w = d.v;
        } else {
            const float u = g.uniform_f32();
            pr = th + delta * (two * u - one);
            if (pr >= pi)
                pr -= two_pi;
            else if (pr < -pi)
                pr += two_pi;
        }
        float sp, cp;
        fast_sincosf(pr, sp, cp);
        const float Cs = ((cu + cd) + cl) + cr;
        const float Ss = ((su + sd) + sl) + sr;
        float t = (cp - c0) * Cs;
        t = fmaf(sp - s0, Ss, t);
        const float dE = -J * t;
        if (MODEL != MODEL_XY_GAUSS) w = g.uniform_f32();
        acc = w < accept_weight(-beta * dE);  // dE <= 0 or w < exp(-beta dE)
        if (acc) {
            v.th[own] = pr;
            v.cs[self] = make_float2(cp, sp);
        }
    }
    ulonglong2 out;
    out.x = g.s0;
    out.y = g.s1;
    v.rng[own] = out;
    return acc;
}

// Observables of the site in LDS row lr (1..H), column j: right and down bonds
// (the row below row H is the ghost row H+1).  Adds the integer energy term
// to e and, when `all`, magnetisation (and XY twist sums) to m.
template <int MODEL>
__device__ __forceinline__ void site_observables(const StripView &v, const int lr, const int j, const int L,
                                                 const bool all, long long &e, double (&m)[4]) {
    const int self = lds_index(lr, j, L), dn = lds_index(lr + 1, j, L);
    const int rt = lds_index(lr, j == L - 1 ? 0 : j + 1, L);
    if (MODEL == MODEL_ISING) {
        const int sv = v.sp[self];
        e += sv * ((int)v.sp[dn] + (int)v.sp[rt]);
        if (all) m[0] += sv;
    } else {
        const float2 o = v.cs[self], d = v.cs[dn], r = v.cs[rt];
        const float cs = o.x, ss = o.y;
        const float bxc = fmaf(ss, d.y, cs * d.x);
        const float byc = fmaf(ss, r.y, cs * r.x);
        e += bond_fixed(bxc + byc);
        if (all) {
            m[0] += cs;
            m[1] += ss;
            m[2] += bxc;
            m[3] += fmaf(ss, d.x, -(cs * d.y));
        }
    }
}

}  // namespace mc
