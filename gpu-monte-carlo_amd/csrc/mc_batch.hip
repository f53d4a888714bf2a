// Persistent batched engine: a whole run (many sweeps, replica exchanges and
// samples) in one cooperative launch.
//
// Decomposition: each replica lattice (L x L) is cut into S = L / H strips of
// H full rows; one 256-thread workgroup owns one strip for the whole launch.
// Everything the strip needs lives in LDS: the RNG streams and state of its
// own rows, and the data neighbours read (cos/sin of the angle, or the int8
// spin) for its rows plus one halo row above and below.  Rows are stored
// colour-split (even columns, then odd columns) so every access of a colour
// step is unit-stride.  Global memory is touched only at launch start/end,
// for the halo rows (two per strip per sweep) and for exchanges and samples.
//
// Per sweep a strip hands two halo rows to its ring neighbours through global
// memory and per-strip flags: its top row after colours 0/1, its bottom row
// after colours 2/3.  Interior rows are updated before each wait (see the
// schedule in the sweep loop), so the hand-off latency overlaps with work.
// Strips of one replica are placed on one XCD.  Every `exchange_interval`
// sweeps the strips publish energy partials and meet at a grid barrier; every
// workgroup then evaluates the same exchange matching on identical inputs
// (identical decisions, no second barrier) and workgroup 0 writes the PT
// state back to global memory.
//
// Every wait is bounded: past a watchdog count a workgroup records an error
// and leaves, so a broken invariant ends the launch instead of hanging it.
//
// Models:
//  * MODEL_ISING: the reference's Ising update, bit-identical to the
//    per-colour kernels (thresholds exp(-beta dE) recomputed per beta change).
//  * MODEL_XY_FAST: the XY update with dE from cached (cos, sin) of every site
//    and the bit-reproducible float32 sincos/exp of mc_fastmath.h.
#include "mc_common.h"
#include "mc_fastmath.h"

namespace mc {

enum { MODEL_ISING = 0, MODEL_XY_FAST = 1 };
enum { BATCH_ERR_WATCHDOG = 1 };

constexpr int kBatchThreads = 256;
constexpr unsigned kWatchdog = 1u << 20;  // about a second of polling
constexpr int kChannels = 5;               // XY: e, mx, my, cx, qx ; Ising: e, m

struct BatchArgs {
    int model, L, R, H, S;
    void *state;                // int8 or float32 (R, L, L)
    uint64_t *rng;              // R*L*L records
    float *betas;               // (R) walker betas, updated at exchanges
    float J, delta;
    int nsweeps;
    long long sweep0;           // global sweep count before this launch
    // parallel tempering (exchange_interval == 0: none)
    int exchange_interval;
    int *parity;                // next parity, in/out
    double *energies;           // (R) walker energies of the last exchange
    int32_t *w2t, *t2w;         // (R)
    uint64_t *pt_rng;           // (R-1) records
    long long *counts;          // (3, R-1)
    long long *trips;           // (5, R)
    // sampling (sample_interval == 0: none)
    int sample_interval;
    long long sample_first;     // global sweep of sample 0
    int max_samples;
    double *sample_partials;    // [max_samples][R][S][kChannels]
    int32_t *sample_w2t;        // [max_samples][R]
    // synchronisation
    float *halo;                // [R][S][2 (top, bottom)][2 channels][L]
    unsigned *flags;            // [R][S][2] (launch-local counters, zeroed by host)
    unsigned *barrier;          // [2] count, generation
    double *epart;              // [2][R][S]
    int *error;
};

// LDS position of column j in LDS row `row` (colour-split layout)
__device__ __forceinline__ int lds_index(int row, int j, int L) {
    return row * L + (j & 1) * (L >> 1) + (j >> 1);
}

// Hand-off protocol.  Every word that crosses workgroups (halo rows, energy
// partials, flags, barrier counters) is written and read with agent-scope
// relaxed atomics, which are device-coherent on their own (gfx950 sets the
// device-scope cache bits on them), so no L2 write-back or invalidate is
// needed:
//  * publisher: every thread stores its part, waits until its stores are
//    complete (s_waitcnt), workgroup barrier, then thread 0 stores the flag;
//  * consumer: thread 0 polls the flag, workgroup barrier, then the data
//    loads are issued -- after the flag was seen, so they read the new data.
struct Sync {
    const BatchArgs &a;
    __device__ unsigned *flag(int rr, int ss, int which) const {
        return a.flags + ((size_t)rr * a.S + ss) * 2 + which;
    }
    // thread 0: poll until *f >= target (or != target), bounded by the watchdog
    __device__ bool poll(unsigned *f, unsigned target, bool until_changed) const {
        unsigned n = 0;
        for (;;) {
            const unsigned v = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (until_changed ? (v != target) : (v >= target)) return true;
            __builtin_amdgcn_s_sleep(1);
            if ((++n & 1023u) == 0 &&
                (n > kWatchdog || __hip_atomic_load(a.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                atomicOr(a.error, BATCH_ERR_WATCHDOG);
                return false;
            }
        }
    }
    __device__ bool wait_flag(unsigned *f, unsigned target) const {
        __shared__ int ok;
        if (threadIdx.x == 0) ok = poll(f, target, false);
        __syncthreads();
        return ok;
    }
    __device__ void release_flag(unsigned *f, unsigned value) const {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (threadIdx.x == 0) __hip_atomic_store(f, value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // sense-reversing grid barrier over all workgroups
    __device__ bool grid_barrier(unsigned nwg) const {
        __shared__ int ok;
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned *count = a.barrier, *gen = a.barrier + 1;
            const unsigned g = __hip_atomic_load(gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned arrived = __hip_atomic_fetch_add(count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            int good = 1;
            if (arrived == nwg - 1) {
                __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __builtin_amdgcn_s_waitcnt(0);
                __hip_atomic_store(gen, g + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                good = poll(gen, g, true);
            }
            ok = good;
        }
        __syncthreads();
        return ok;
    }
};

__device__ __forceinline__ void store_relaxed(float *p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float load_relaxed(const float *p) {
    return __hip_atomic_load(const_cast<float *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Ising acceptance threshold for 2 s sum_nn = v: always flip when dE <= 0,
// else u < exp(-beta dE) in float32 through the float64 exponential.
__device__ __forceinline__ float ising_threshold(int v, float beta, float J) {
    const float dE = (float)v * J;
    if (dE <= 0.0f) return __builtin_inff();
    return rexp(-beta * dE);
}

// Block reduction of C float64 channels; the result is valid in thread 0.
template <int C>
__device__ void block_sum(double (&v)[C], double (*sh)[kBatchThreads / 64]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < C; c++) {
        double x = v[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if (lane == 0) sh[c][wave] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < C; c++) v[c] = ((sh[c][0] + sh[c][1]) + sh[c][2]) + sh[c][3];
    }
    __syncthreads();
}

// LDS bytes of one strip's site data (everything but the PT scratch)
__host__ __device__ inline size_t strip_bytes(int model, int L, int H) {
    const size_t rng = (size_t)H * L * 16;
    return model == MODEL_ISING ? rng + (size_t)(H + 2) * L
                                : rng + (size_t)(H + 2) * L * 8 + (size_t)H * L * 4;
}

template <int MODEL>
__global__ __launch_bounds__(kBatchThreads) void batch_kernel(BatchArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int L = a.L, H = a.H, S = a.S, R = a.R;
    const int half = L >> 1, HH = H >> 1;
    const int tid = threadIdx.x;
    const unsigned nwg = gridDim.x;

    // XCD-aware placement: all strips of a replica on one XCD when possible
    int lin = blockIdx.x;
    if (nwg % 8 == 0 && (nwg / 8) % S == 0) lin = (blockIdx.x % 8) * (nwg / 8) + blockIdx.x / 8;
    const int r = lin / S, s = lin % S;
    const Sync sync{a};
    const int s_up = (s + S - 1) % S, s_dn = (s + 1) % S;

    // ---- LDS carve-up
    ulonglong2 *lrng = reinterpret_cast<ulonglong2 *>(smem);            // own rows, H * L
    float *lc = reinterpret_cast<float *>(lrng + (size_t)H * L);         // rows 0..H+1
    float *ls = lc + (H + 2) * L;                                        // rows 0..H+1
    float *lth = ls + (H + 2) * L;                                       // own rows, H * L
    int8_t *lsp = reinterpret_cast<int8_t *>(lrng + (size_t)H * L);      // rows 0..H+1 (Ising)
    unsigned char *pt_base = smem + ((strip_bytes(MODEL, L, H) + 15) & ~(size_t)15);
    float *pt_beta = reinterpret_cast<float *>(pt_base);                 // R
    int32_t *pt_t2w = reinterpret_cast<int32_t *>(pt_beta + R);          // R
    int32_t *pt_w2t = pt_t2w + R;                                         // R
    double *pt_e = reinterpret_cast<double *>(pt_base + ((size_t)(3 * R * 4 + 7) & ~(size_t)7));  // R
    uint64_t *pt_g = reinterpret_cast<uint64_t *>(pt_e + R);             // 2 (R-1)
    __shared__ double red[kChannels][kBatchThreads / 64];
    __shared__ float thr[5];
    __shared__ float my_beta;

    const size_t rep = (size_t)r * L * L;
    const int row0 = s * H;  // first global row of the strip
    const bool use_pt = a.exchange_interval > 0;

    // ---- load the strip: own rows (state, RNG) and neighbour data incl. halos
    for (int idx = tid; idx < (H + 2) * L; idx += kBatchThreads) {
        const int lr = idx / L, j = idx - lr * L;
        const int gi = (row0 + lr - 1 + L) % L;
        const size_t gk = rep + (size_t)gi * L + j;
        const bool own_row = lr >= 1 && lr <= H;
        if (own_row) lrng[lds_index(lr - 1, j, L)] = reinterpret_cast<const ulonglong2 *>(a.rng)[gk];
        if (MODEL == MODEL_ISING) {
            lsp[lds_index(lr, j, L)] = reinterpret_cast<const int8_t *>(a.state)[gk];
        } else {
            const float th = reinterpret_cast<const float *>(a.state)[gk];
            float sv, cv;
            fast_sincosf(th, sv, cv);
            lc[lds_index(lr, j, L)] = cv;
            ls[lds_index(lr, j, L)] = sv;
            if (own_row) lth[lds_index(lr - 1, j, L)] = th;
        }
    }
    if (use_pt) {
        for (int k = tid; k < R; k += kBatchThreads) {
            pt_beta[k] = a.betas[k];
            pt_t2w[k] = a.t2w[k];
            pt_w2t[k] = a.w2t[k];
        }
        for (int k = tid; k < 2 * (R - 1); k += kBatchThreads) pt_g[k] = a.pt_rng[k];
    }
    if (tid == 0) my_beta = a.betas[r];
    __syncthreads();
    if (MODEL == MODEL_ISING && tid < 5) thr[tid] = ising_threshold(4 * tid - 8, my_beta, a.J);
    int parity = use_pt ? *a.parity : 0;
    __syncthreads();

    const float J = a.J, delta = a.delta;
    const float one = 1.0f, two = 2.0f;
    const float pi = 3.14159274101257324f, two_pi = 6.28318548202514648f;
    // fixed thread -> (column, row offset) map of a colour step
    const int tcol = tid % half, trow = tid / half, rows_per_pass = kBatchThreads / half;

    // Update the sites of colour c in colour rows [ri_lo, ri_hi]; colour rows of
    // colours 0/1 are local rows 1, 3, ...; of colours 2/3 local rows 2, 4, ...
    auto update = [&](const int c, const int ri_lo, const int ri_hi, const float beta) {
        if (trow >= rows_per_pass) return;
        const int p = c & 1;
        for (int ri = ri_lo + trow; ri <= ri_hi; ri += rows_per_pass) {
            const int lr = 1 + 2 * ri + (c >> 1);
            const int col = tcol;
            const int base = lr * L;
            const int self = base + p * half + col;
            const int up = self - L, dn = self + L;
            const int lf = p ? base + col : base + half + (col == 0 ? half - 1 : col - 1);
            const int rt = p ? base + (col == half - 1 ? 0 : col + 1) : base + half + col;
            const int own = self - L;  // own-row arrays start at local row 1
            const ulonglong2 st = lrng[own];
            Rng g;
            g.s0 = st.x;
            g.s1 = st.y;
            if (MODEL == MODEL_ISING) {
                const int sv = lsp[self];
                const int nb = (int)lsp[up] + (int)lsp[dn] + (int)lsp[lf] + (int)lsp[rt];
                const float u = g.uniform_f32();
                if (u < thr[(2 * sv * nb + 8) >> 2]) lsp[self] = (int8_t)(-sv);
            } else {
                const float th = lth[own];
                const float u = g.uniform_f32();
                float pr = th + delta * (two * u - one);
                if (pr >= pi)
                    pr -= two_pi;
                else if (pr < -pi)
                    pr += two_pi;
                float sp, cp;
                fast_sincosf(pr, sp, cp);
                const float Cs = ((lc[up] + lc[dn]) + lc[lf]) + lc[rt];
                const float Ss = ((ls[up] + ls[dn]) + ls[lf]) + ls[rt];
                float t = (cp - lc[self]) * Cs;
                t = fmaf(sp - ls[self], Ss, t);
                const float dE = -J * t;
                const float v = g.uniform_f32();
                if (dE <= 0.0f || v < fast_expf(-beta * dE)) {
                    lth[own] = pr;
                    lc[self] = cp;
                    ls[self] = sp;
                }
            }
            ulonglong2 out;
            out.x = g.s0;
            out.y = g.s1;
            lrng[own] = out;
        }
    };
    auto halo_ptr = [&](int ss, int which) { return a.halo + (((size_t)r * S + ss) * 2 + which) * 2 * L; };
    auto load_row = [&](int lds_row, const float *src) {
        for (int j = tid; j < L; j += kBatchThreads) {
            if (MODEL == MODEL_ISING) {
                lsp[lds_index(lds_row, j, L)] = (int8_t)load_relaxed(src + j);
            } else {
                lc[lds_index(lds_row, j, L)] = load_relaxed(src + j);
                ls[lds_index(lds_row, j, L)] = load_relaxed(src + L + j);
            }
        }
    };
    auto store_row = [&](float *dst, int lds_row) {
        for (int j = tid; j < L; j += kBatchThreads) {
            if (MODEL == MODEL_ISING) {
                store_relaxed(dst + j, (float)lsp[lds_index(lds_row, j, L)]);
            } else {
                store_relaxed(dst + j, lc[lds_index(lds_row, j, L)]);
                store_relaxed(dst + L + j, ls[lds_index(lds_row, j, L)]);
            }
        }
    };
    auto copy_row = [&](int dst_row, int src_row) {
        for (int j = tid; j < L; j += kBatchThreads) {
            if (MODEL == MODEL_ISING) {
                lsp[lds_index(dst_row, j, L)] = lsp[lds_index(src_row, j, L)];
            } else {
                lc[lds_index(dst_row, j, L)] = lc[lds_index(src_row, j, L)];
                ls[lds_index(dst_row, j, L)] = ls[lds_index(src_row, j, L)];
            }
        }
    };

    // One sweep = colours 0, 1, 2, 3 in the reference order.  A site's update
    // reads only neighbouring colours, so any order in which
    //   c1(row) follows c0(row), c2(row) follows c0(rows +-1),
    //   c3(row) follows c2(row) and c1(rows +-1),
    //   and each site's update precedes its neighbours' next update
    // gives identical results.  Interior rows therefore go first: only the top
    // row T (local 1) and T + 1 depend on the strip above (its bottom row after
    // colours 2/3 of the previous sweep), and only the bottom row B (local H)
    // on the strip below (its top row after colours 0/1 of this sweep).
    for (int n = 0; n < a.nsweeps; n++) {
        const long long sweep_global = a.sweep0 + n + 1;
        const float beta = my_beta;
        update(0, 1, HH - 1, beta);                  // c0, interior rows of colours 0/1
        __syncthreads();
        update(1, 1, HH - 1, beta);                  // c1, same rows
        update(2, 1, HH - 2, beta);                  // c2, interior rows of colours 2/3
        __syncthreads();
        if (n > 0 && S > 1) {                        // halo above: row B of strip s-1
            if (!sync.wait_flag(sync.flag(r, s_up, 1), (unsigned)n)) goto done;
            load_row(0, halo_ptr(s_up, 1));
            __syncthreads();
        }
        update(3, 1, HH - 2, beta);                  // c3, interior rows of colours 2/3
        update(0, 0, 0, beta);                       // c0, row T
        __syncthreads();
        update(1, 0, 0, beta);                       // c1, row T
        update(2, 0, 0, beta);                       // c2, row T + 1
        __syncthreads();
        if (S > 1) {                                 // publish row T for strip s-1
            store_row(halo_ptr(s, 0), 1);
            sync.release_flag(sync.flag(r, s, 0), (unsigned)(n + 1));
        } else {
            copy_row(H + 1, 1);                      // one strip: the row below B is T
        }
        update(3, 0, 0, beta);                       // c3, row T + 1
        __syncthreads();
        if (S > 1) {                                 // halo below: row T of strip s+1
            if (!sync.wait_flag(sync.flag(r, s_dn, 0), (unsigned)(n + 1))) goto done;
            load_row(H + 1, halo_ptr(s_dn, 0));
            __syncthreads();
        }
        update(2, HH - 1, HH - 1, beta);             // c2, row B
        __syncthreads();
        update(3, HH - 1, HH - 1, beta);             // c3, row B
        __syncthreads();
        if (S > 1) {                                 // publish row B for strip s+1
            store_row(halo_ptr(s, 1), H);
            sync.release_flag(sync.flag(r, s, 1), (unsigned)(n + 1));
        } else {
            copy_row(0, H);                          // one strip: the row above T is B
            __syncthreads();
        }

        const bool do_exchange = use_pt && (sweep_global % a.exchange_interval == 0);
        const bool do_sample = a.sample_interval > 0 && sweep_global >= a.sample_first &&
                               ((sweep_global - a.sample_first) % a.sample_interval == 0);
        if (!do_exchange && !do_sample) continue;

        // ---- strip observables: right and down bonds (row below = LDS row H+1,
        // current: it only changes in colours 0/1)
        double v[kChannels];
        {
            double e = 0, m1 = 0, m2 = 0, cx = 0, qx = 0;
            float ek = 0.0f, ec = 0.0f;  // compensated sum of the XY bond energies
            for (int idx = tid; idx < H * L; idx += kBatchThreads) {
                const int lr = 1 + idx / L, j = idx % L;
                const int self = lds_index(lr, j, L), dn = lds_index(lr + 1, j, L);
                const int rt = lds_index(lr, j == L - 1 ? 0 : j + 1, L);
                if (MODEL == MODEL_ISING) {
                    const int sv = lsp[self];
                    e += -(double)J * sv * ((int)lsp[dn] + (int)lsp[rt]);
                    m1 += sv;
                } else {
                    const float cs = lc[self], ss = ls[self];
                    const float bxc = fmaf(ss, ls[dn], cs * lc[dn]);
                    const float byc = fmaf(ss, ls[rt], cs * lc[rt]);
                    const float bxs = fmaf(ss, lc[dn], -(cs * ls[dn]));
                    const float y = -J * (bxc + byc) - ec;
                    const float t = ek + y;
                    ec = (t - ek) - y;
                    ek = t;
                    m1 += cs;
                    m2 += ss;
                    cx += bxc;
                    qx += bxs;
                }
            }
            if (MODEL == MODEL_XY_FAST) e = (double)ek - (double)ec;
            v[0] = e;
            v[1] = m1;
            v[2] = m2;
            v[3] = cx;
            v[4] = qx;
        }
        block_sum<kChannels>(v, red);

        if (do_exchange) {
            const int buf = (int)((sweep_global / a.exchange_interval) & 1);
            double *ep = a.epart + ((size_t)buf * R + r) * S + s;
            if (tid == 0) __hip_atomic_store(ep, v[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (!sync.grid_barrier(nwg)) goto done;
            // every workgroup: the same total energies, in the same order
            for (int k = tid; k < R; k += kBatchThreads) {
                double tot = 0.0;
                const double *src = a.epart + ((size_t)buf * R + k) * S;
                for (int q = 0; q < S; q++)
                    tot += __hip_atomic_load(const_cast<double *>(src + q), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
                pt_e[k] = tot;
            }
            __syncthreads();
            // one matching of parity `parity` (core/parallel_tempering.py:44-69)
            const int P = R - 1;
            for (int left = parity + 2 * tid; left < R - 1; left += 2 * kBatchThreads) {
                const int right = left + 1;
                const int wa = pt_t2w[left], wb = pt_t2w[right];
                const double ea = pt_e[wa], eb = pt_e[wb];
                const bool finite = isfinite(ea) && isfinite(eb);
                bool accept = false;
                if (finite) {
                    const double lr = ((double)pt_beta[wa] - (double)pt_beta[wb]) * (ea - eb);
                    Rng g;
                    g.s0 = pt_g[2 * left];
                    g.s1 = pt_g[2 * left + 1];
                    const double u = g.uniform_f64();
                    pt_g[2 * left] = g.s0;
                    pt_g[2 * left + 1] = g.s1;
                    accept = lr >= 0.0 || u == 0.0 || ::log(u) < lr;
                }
                if (accept) {
                    pt_t2w[left] = wb;
                    pt_t2w[right] = wa;
                    pt_w2t[wa] = right;
                    pt_w2t[wb] = left;
                    const float tb = pt_beta[wa];
                    pt_beta[wa] = pt_beta[wb];
                    pt_beta[wb] = tb;
                }
                if (blockIdx.x == 0) {
                    a.counts[left] += 1;
                    if (!finite) a.counts[2 * P + left] += 1;
                    if (accept) a.counts[P + left] += 1;
                }
            }
            __syncthreads();
            if (blockIdx.x == 0) {
                // round-trip bookkeeping (core/parallel_tempering.py:71-91)
                for (int w = tid; w < R; w += kBatchThreads) {
                    const int t = pt_w2t[w];
                    if (t != 0 && t != R - 1) continue;
                    const long long origin = a.trips[w];
                    if (origin < 0) {
                        a.trips[w] = t;
                        a.trips[2 * R + w] = sweep_global;
                    } else if (t != origin) {
                        a.trips[R + w] = 1;
                    } else if (a.trips[R + w]) {
                        a.trips[3 * R + w] += 1;
                        a.trips[4 * R + w] += sweep_global - a.trips[2 * R + w];
                        a.trips[2 * R + w] = sweep_global;
                        a.trips[R + w] = 0;
                    }
                }
                for (int k = tid; k < R; k += kBatchThreads) {
                    a.betas[k] = pt_beta[k];
                    a.t2w[k] = pt_t2w[k];
                    a.w2t[k] = pt_w2t[k];
                    a.energies[k] = pt_e[k];
                }
                for (int k = tid; k < 2 * (R - 1); k += kBatchThreads) a.pt_rng[k] = pt_g[k];
            }
            if (R > 2) parity = 1 - parity;
            if (tid == 0) my_beta = pt_beta[r];
            __syncthreads();
            if (MODEL == MODEL_ISING && tid < 5) thr[tid] = ising_threshold(4 * tid - 8, my_beta, a.J);
            __syncthreads();
        }

        if (do_sample) {
            const long long idx = (sweep_global - a.sample_first) / a.sample_interval;
            if (idx < a.max_samples) {
                if (tid == 0) {
                    double *dst = a.sample_partials + (((size_t)idx * R + r) * S + s) * kChannels;
#pragma unroll
                    for (int c = 0; c < kChannels; c++) dst[c] = v[c];
                }
                if (blockIdx.x == 0)
                    for (int k = tid; k < R; k += kBatchThreads)
                        a.sample_w2t[(size_t)idx * R + k] = use_pt ? pt_w2t[k] : k;
            }
        }
    }

done:
    // ---- write back own rows: state and RNG streams
    __syncthreads();
    for (int idx = tid; idx < H * L; idx += kBatchThreads) {
        const int lr = 1 + idx / L, j = idx - (lr - 1) * L;
        const size_t gk = rep + (size_t)(row0 + lr - 1) * L + j;
        reinterpret_cast<ulonglong2 *>(a.rng)[gk] = lrng[lds_index(lr - 1, j, L)];
        if (MODEL == MODEL_ISING)
            reinterpret_cast<int8_t *>(a.state)[gk] = lsp[lds_index(lr, j, L)];
        else
            reinterpret_cast<float *>(a.state)[gk] = lth[lds_index(lr - 1, j, L)];
    }
    if (blockIdx.x == 0 && tid == 0 && use_pt) *a.parity = parity;
}

static size_t batch_lds(int model, int L, int R, int H, int exchange) {
    const size_t pt_bytes = exchange ? (size_t)3 * R * 4 + 8 + (size_t)R * 8 + (size_t)2 * R * 8 : 0;
    return ((strip_bytes(model, L, H) + 15) & ~(size_t)15) + pt_bytes + 64;
}

static const void *kernel_ptr(int model) {
    return model == MODEL_ISING ? reinterpret_cast<const void *>(&batch_kernel<MODEL_ISING>)
                                : reinterpret_cast<const void *>(&batch_kernel<MODEL_XY_FAST>);
}

static int check_shape(int model, int L, int R, int H) {
    MC_REQUIRE(model == MODEL_ISING || model == MODEL_XY_FAST, "unknown batch model %d", model);
    MC_REQUIRE(L >= 4 && L % 2 == 0, "L must be even and >= 4, got %d", L);
    MC_REQUIRE(H >= 4 && H % 2 == 0 && L % H == 0, "strip height %d must be even, >= 4 and divide L=%d",
               H, L);
    MC_REQUIRE(L / 2 <= kBatchThreads, "L=%d is too wide for one workgroup (L/2 <= %d)", L, kBatchThreads);
    MC_REQUIRE(R >= 1, "need at least one replica");
    return MC_OK;
}

}  // namespace mc

using namespace mc;

extern "C" {

int mc_batch_workspace_size(int model, int L, int R, int H, size_t *halo_floats, size_t *flag_words,
                            size_t *epart_doubles) {
    int rc = check_shape(model, L, R, H);
    if (rc) return rc;
    const int S = L / H;
    *halo_floats = (size_t)R * S * 2 * 2 * L;
    *flag_words = (size_t)R * S * 2 + 2;
    *epart_doubles = (size_t)2 * R * S;
    return MC_OK;
}

int mc_batch_max_workgroups(int model, int L, int R, int H, int exchange, int *max_workgroups) {
    int rc = check_shape(model, L, R, H);
    if (rc) return rc;
    const size_t lds = batch_lds(model, L, R, H, exchange);
    if (lds > 160 * 1024) {
        *max_workgroups = 0;
        return MC_OK;
    }
    if (lds > 64 * 1024)
        MC_TRY_HIP(hipFuncSetAttribute(kernel_ptr(model), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0, dev = 0, cus = 0;
    MC_TRY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_ptr(model), kBatchThreads, lds));
    MC_TRY_HIP(hipGetDevice(&dev));
    MC_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    *max_workgroups = per_cu * cus;
    return MC_OK;
}

int mc_batch_run(void *stream, int model, int L, int R, int H, void *state, uint64_t *rng, float *betas,
                 float J, float delta, int nsweeps, long long sweep0, int exchange_interval, int *parity,
                 double *energies, int32_t *w2t, int32_t *t2w, uint64_t *pt_rng, long long *counts,
                 long long *trips, int sample_interval, long long sample_first, int max_samples,
                 double *sample_partials, int32_t *sample_w2t, float *halo, unsigned *flags,
                 double *epart, int *error) {
    int rc = check_shape(model, L, R, H);
    if (rc) return rc;
    MC_REQUIRE(nsweeps >= 0, "nsweeps must be nonnegative");
    MC_REQUIRE(exchange_interval >= 0, "exchange_interval must be nonnegative");
    MC_REQUIRE(exchange_interval == 0 || R >= 2, "exchanges need at least two replicas");
    MC_REQUIRE(sample_interval >= 0 && max_samples >= 0, "invalid sampling arguments");
    const int S = L / H;
    const size_t lds = batch_lds(model, L, R, H, exchange_interval > 0);
    MC_REQUIRE(lds <= 160 * 1024, "batch workgroup needs %zu bytes of LDS", lds);
    if (nsweeps == 0) return MC_OK;

    BatchArgs a;
    a.model = model; a.L = L; a.R = R; a.H = H; a.S = S;
    a.state = state; a.rng = rng; a.betas = betas; a.J = J; a.delta = delta;
    a.nsweeps = nsweeps; a.sweep0 = sweep0;
    a.exchange_interval = exchange_interval; a.parity = parity; a.energies = energies;
    a.w2t = w2t; a.t2w = t2w; a.pt_rng = pt_rng; a.counts = counts; a.trips = trips;
    a.sample_interval = sample_interval; a.sample_first = sample_first; a.max_samples = max_samples;
    a.sample_partials = sample_partials; a.sample_w2t = sample_w2t;
    a.halo = halo; a.flags = flags; a.barrier = flags + (size_t)R * S * 2; a.epart = epart; a.error = error;

    if (lds > 64 * 1024)
        MC_TRY_HIP(hipFuncSetAttribute(kernel_ptr(model), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // launch-local flag counters and barrier start at zero
    MC_TRY_HIP(hipMemsetAsync(flags, 0, ((size_t)R * S * 2 + 2) * sizeof(unsigned), as_stream(stream)));
    MC_TRY_HIP(hipMemsetAsync(error, 0, sizeof(int), as_stream(stream)));
    void *kargs[] = {&a};
    MC_TRY_HIP(hipLaunchCooperativeKernel(kernel_ptr(model), dim3((unsigned)(R * S)), dim3(kBatchThreads),
                                          kargs, lds, as_stream(stream)));
    return MC_OK;
}

}  // extern "C"
