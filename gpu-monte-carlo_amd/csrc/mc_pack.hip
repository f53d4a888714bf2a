// Packed persistent engine: the runs of several lattice sizes (and models) --
// a finite-size-scaling set -- in one launch.
//
// A group is one (model, L, ladder) ensemble of R replicas; every group keeps
// the reference's schedule of sweeps, exchanges every `exchange_interval`
// sweeps (core/parallel_tempering.py:245-269) and samples.  A group's
// replicas are cut into strips of H rows (mc_strip.h), and a task is one
// strip position s of a block of K consecutive replicas.  Each workgroup runs
// a fixed list of tasks of one group, all of them every sweep: the sizes of a
// whole study do not fit on chip together (the RNG streams alone are 16 bytes
// per site), so a task is streamed through LDS each sweep -- rows 1..H from
// the lattice arrays, the ghost row from its own buffer, rows 0 and H+2 from
// the neighbours' halo words, loaded before the strip so that their latency
// overlaps it -- then swept colour by colour in the reference order and
// written back.  One sweep of a task reads and writes each of its sites' RNG
// stream and state once, plus one ghost row and two halo rows per strip.
//
// Ordering.  A task at sweep n needs only its neighbours' rows as of the end
// of sweep n-1, published as tagged halo words (tag n+1 = "at the start of
// sweep n", slot n & 1), so any placement of tasks on co-resident workgroups
// makes progress: the dependencies of (sweep n, task i of a workgroup) are
// tasks of sweep n-1 and earlier tasks of the same workgroup.  The launch
// must be fully co-resident (checked on the host against the occupancy of
// its LDS size); every wait is bounded by a watchdog regardless.
//
// Exchanges.  Energies are integers (mc_strip.h).  Each task publishes the
// strip energy of each of its replicas as two tagged words (low and high
// halves, tag = global exchange index + 1, double-buffered by exchange
// parity); every workgroup of the group polls all R * S word pairs, adds each
// replica's strips as integers and runs the identical matching
// (core/parallel_tempering.py:44-91) on identical inputs.  A word carries its
// own tag, so no fence or cache maintenance is needed (an acquire/release at
// device scope writes back and invalidates the L2 on gfx950).  A workgroup
// cannot publish exchange x + 2 before every workgroup of its group has
// published exchange x + 1, which each does only after reading exchange x,
// so two buffers suffice.  Counters of pair `left` and trip records of walker w are kept by
// the group's workgroup left (w) mod its workgroup count; the group's first
// workgroup writes the PT state back at the end of the launch.
//
// Graphs.  The sweep count lives in device memory (`clock`), so a captured
// prologue + launch can be replayed: each replay continues the chain.  The
// prologue (a kernel, not memset nodes) refreshes the ghost rows, publishes
// the halo words for sweep 0 and clears the exchange totals.
#include <algorithm>
#include <vector>

#include "mc_strip.h"

namespace mc {

// Pointers read from the group table are generic; these name them as global
// memory so that loads and stores compile to global (not flat) instructions.
template <class T>
using gmem = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gmem<T> as_global(void *p) {
    return (gmem<T>)(T *)p;
}
template <class T>
__device__ __forceinline__ gmem<const T> as_global(const void *p) {
    return (gmem<const T>)(const T *)p;
}
// one RNG record in global memory (member-wise access, merged into one
// 16-byte load or store)
struct alignas(16) RngRec {
    unsigned long long s0, s1;
};
__device__ __forceinline__ ulonglong2 load_rec(gmem<RngRec> p) {
    ulonglong2 r;
    r.x = p->s0;
    r.y = p->s1;
    return r;
}
__device__ __forceinline__ void store_rec(gmem<RngRec> p, const ulonglong2 &r) {
    p->s0 = r.x;
    p->s1 = r.y;
}

struct PackGroup {
    int model, L, R, H, S, K;
    int ntasks;                // (R / K) * S; task t = (replica block t / S, strip t % S)
    int wg_first, wg_count;    // the group's workgroups
    int exchange_interval;     // 0: no exchanges
    float J, delta;
    void *state;               // int8 or float32 (R, L, L)
    uint64_t *rng;             // R*L*L records
    float *betas;              // (R) walker betas
    const float *sigmas;       // (R) per temperature, MODEL_XY_GAUSS
    unsigned long long *accepted;  // (R) per temperature, may be NULL
    int *parity;
    double *energies;          // (R)
    int32_t *w2t, *t2w;
    uint64_t *pt_rng;          // (R-1) records
    long long *counts;         // (3, R-1)
    long long *trips;          // (5, R)
    // workspace owned by the pack
    uint64_t *halo;            // [R][S][2 (row 2, row H)][2 (slot)][L] words
    uint64_t *ghost_rng;       // [R][S][L] records
    void *ghost_state;         // [R][S][L]
    uint64_t *energy_words;    // [2 (exchange parity)][R][S][2 (lo, hi)] tagged words
};

// A value the same in every lane, moved to scalar registers
template <class T>
__device__ __forceinline__ T uniform(T v) {
    static_assert(sizeof(T) == 4 || sizeof(T) == 8, "32- or 64-bit values");
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
    } else {
        const unsigned long long x = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)x);
        const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(x >> 32));
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    }
}

// The fields of the descriptor that the sweep loop uses, read once at entry
// into scalar registers (the others -- PT state, counters, records -- are
// read through the table where they are used, at entry, exchanges and exit).
// Read through the table where each is used, the loads inside the loop could
// be neither scalar nor hoisted (the loop's stores might alias the table):
// each became a vector load waited for on the spot, several memory round
// trips per streamed record.
struct PackHot {
    int L, R, H, S, K, wg_count, exchange_interval;
    float J, delta;
    void *state;
    uint64_t *rng;
    const float *sigmas;
    uint64_t *halo, *ghost_rng;
    void *ghost_state;
    uint64_t *energy_words;
};
__device__ __forceinline__ PackHot load_hot(const PackGroup *gp) {
    PackHot g;
    g.L = uniform(gp->L); g.R = uniform(gp->R); g.H = uniform(gp->H); g.S = uniform(gp->S); g.K = uniform(gp->K);
    g.wg_count = uniform(gp->wg_count); g.exchange_interval = uniform(gp->exchange_interval);
    g.J = uniform(gp->J); g.delta = uniform(gp->delta);
    g.state = uniform(gp->state); g.rng = uniform(gp->rng); g.sigmas = uniform(gp->sigmas);
    g.halo = uniform(gp->halo); g.ghost_rng = uniform(gp->ghost_rng); g.ghost_state = uniform(gp->ghost_state);
    g.energy_words = uniform(gp->energy_words);
    return g;
}

struct PackWork {
    int group, task_first, task_count, pad;
};

// Sample buffers travel by value in the launch arguments, so a captured
// graph keeps the buffers it was captured with, whatever later launches use.
constexpr int kMaxPackGroups = 64;

struct PackArgs {
    const PackGroup *groups;
    const PackWork *work;
    int ngroups;
    int nsweeps;
    long long *clock;          // [0] sweeps done, [1] sweeps done before this launch
    int sample_interval;
    long long sample_first;
    int max_samples;
    int *error;
    double *sample_partials[kMaxPackGroups];  // per group [max_samples][R][S][kChannels]
    int32_t *sample_w2t[kMaxPackGroups];      // per group [max_samples][R]
};

// Dynamic LDS of a workgroup of a group: K strips, then the group's PT state
struct PackLds {
    unsigned stride, beta, w2t, t2w, thr, sig, moves, cnt, e, g, sum, trips, total;
};

__host__ __device__ inline PackLds pack_lds(int model, int L, int R, int H, int K) {
    PackLds o;
    o.stride = (unsigned)((strip_bytes(model, L, H) + 15) & ~(size_t)15);
    unsigned p = (unsigned)K * o.stride;
    o.beta = p;
    p += 4u * R;
    o.w2t = p;
    p += 4u * R;
    o.t2w = p;
    p += 4u * R;
    o.thr = p;
    p += 20u * K;
    o.sig = p;
    p += 4u * K;
    o.moves = p;
    p += 4u * R;
    o.cnt = p;
    p += 12u * (R > 1 ? R - 1 : 1);
    p = (p + 7) & ~7u;
    o.e = p;
    p += 8u * R;
    o.g = p;
    p += 16u * (R > 1 ? R - 1 : 1);
    o.sum = p;
    p += 8u * R;
    o.trips = p;
    p += 40u * R;
    o.total = p;
    return o;
}

template <int MODEL, int NT>
__device__ void pack_run(const PackArgs &a, const PackGroup *gp, const PackWork &w, unsigned char *smem,
                         long long *red_e, double (*red)[NT / 16], const float4 *q_tab, int &s_err) {
    const PackHot g = load_hot(gp);
    const PackGroup &gc = *gp;  // the fields used outside the sweep loop's hot path
    const int L = g.L, H = g.H, S = g.S, R = g.R, K = g.K;
    const int HH = H >> 1, half = L >> 1;
    const int tid = threadIdx.x;
    const int lw = (int)blockIdx.x - gc.wg_first;  // workgroup index within the group
    const bool use_pt = g.exchange_interval > 0;
    const float J = g.J, delta = g.delta;

    const PackLds lay = pack_lds(MODEL, L, R, H, K);
    float *pt_beta = reinterpret_cast<float *>(smem + lay.beta);
    int32_t *pt_w2t = reinterpret_cast<int32_t *>(smem + lay.w2t);
    int32_t *pt_t2w = reinterpret_cast<int32_t *>(smem + lay.t2w);
    float *thr = reinterpret_cast<float *>(smem + lay.thr);      // [K][5]
    float *sig = reinterpret_cast<float *>(smem + lay.sig);      // [K]
    unsigned *moves = reinterpret_cast<unsigned *>(smem + lay.moves);  // [R] accepted since the last flush
    double *pt_e = reinterpret_cast<double *>(smem + lay.e);
    uint64_t *pt_g = reinterpret_cast<uint64_t *>(smem + lay.g);
    long long *psum = reinterpret_cast<long long *>(smem + lay.sum);  // [R] integer energies of an exchange
    // round-trip records and swap counters of the walkers and pairs this
    // workgroup keeps (index mod the group's workgroups), in LDS for the launch
    long long *pt_trips = reinterpret_cast<long long *>(smem + lay.trips);  // [5][R]
    unsigned *pt_cnt = reinterpret_cast<unsigned *>(smem + lay.cnt);       // [3][R-1]

    for (int k = tid; k < R; k += NT) {
        pt_beta[k] = gc.betas[k];
        pt_w2t[k] = use_pt ? gc.w2t[k] : k;
        pt_t2w[k] = use_pt ? gc.t2w[k] : k;
        moves[k] = 0;
    }
    if (use_pt) {
        for (int k = tid; k < 2 * (R - 1); k += NT) pt_g[k] = gc.pt_rng[k];
        for (int k = tid; k < 5 * R; k += NT) pt_trips[k] = gc.trips[k];
        for (int k = tid; k < 3 * (R - 1); k += NT) pt_cnt[k] = 0;
    }
    int parity = use_pt ? *gc.parity : 0;
    const long long sweep0 = a.clock[1];
    bool exchanged = false;
    // A wait that expired sets s_err before a barrier; every thread reads it
    // after the barrier (one barrier, where a block-wide OR takes three).
    if (tid == 0) s_err = 0;
    __syncthreads();

    // pass map: colour rows of the K strips spread over the threads
    const int rows_all = NT / half, rpp = rows_all / K;
    const int tcol = tid % half, trow = tid / half;
    const int pk = trow / rpp, prow = trow - pk * rpp;
    const bool pass_on = pk < K;
    // row map of loads, stores, halo rows: whole rows of L sites
    const int lj = tid % L, lrow = tid / L, lrows = NT / L;
    const int ljc = (lj & 1) * half + (lj >> 1);  // column lj in a colour-split row
    // observables map: NT / K threads per replica
    const int TPR = NT / K, ok = tid / TPR, ot = tid - ok * TPR;
    const int W = TPR < 64 ? TPR : 64;

    auto view_of = [&](int k) { return strip_view(smem + (unsigned)k * lay.stride, L, H); };
    auto halo_ptr = [&](int r, int ss, int which, int slot) {
        return g.halo + ((((size_t)r * S + ss) * 2 + which) * 2 + slot) * L;
    };
    auto flush_moves = [&]() {
        __syncthreads();
        if (gc.accepted)
            for (int k = tid; k < R; k += NT)
                if (moves[k]) {
                    atomicAdd(&gc.accepted[pt_w2t[k]], (unsigned long long)moves[k]);
                    moves[k] = 0;
                }
        __syncthreads();
    };

    bool failed = false;
    for (int n = 0; n < a.nsweeps && !failed; n++) {
        const long long sweep_global = sweep0 + n + 1;
        const bool do_exchange = use_pt && (sweep_global % g.exchange_interval == 0);
        const bool do_sample = a.sample_interval > 0 && sweep_global >= a.sample_first &&
                               ((sweep_global - a.sample_first) % a.sample_interval == 0);
        const long long sidx = do_sample ? (sweep_global - a.sample_first) / a.sample_interval : 0;

        for (int ti = 0; ti < w.task_count; ti++) {
            const int t = w.task_first + ti;
            const int b = t / S, s = t - (t / S) * S;
            const int r0 = b * K;
            const int row0 = s * H;
            const int s_up = (s + S - 1) % S, s_dn = (s + 1) % S;
            // ---- prefetch the halo words of rows 0 and H+2 (as of the start of
            // sweep n): their latency overlaps the strip load.  K * L <= NT,
            // so a thread holds at most two of the 2K halo rows.
            const int hq0 = lrow, hq1 = lrow + lrows;
            const bool h0 = S > 1 && lrow < lrows && hq0 < 2 * K;
            const bool h1 = S > 1 && lrow < lrows && hq1 < 2 * K;
            auto halo_src = [&](int qr) {
                const int k = qr >> 1;
                return (qr & 1) ? halo_ptr(r0 + k, s_dn, 0, n & 1) + lj : halo_ptr(r0 + k, s_up, 1, n & 1) + lj;
            };
            uint64_t hw0 = 0, hw1 = 0;
            if (h0) hw0 = __hip_atomic_load(halo_src(hq0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (h1) hw1 = __hip_atomic_load(halo_src(hq1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (tid < K) {
                const int r = r0 + tid;
                if (MODEL == MODEL_ISING)
                    for (int m = 0; m < 5; m++) thr[tid * 5 + m] = ising_threshold(4 * m - 8, pt_beta[r], J);
                if (MODEL == MODEL_XY_GAUSS) sig[tid] = g.sigmas[pt_w2t[r]];
            }
            // ---- load rows 1..H (lattice) and H+1 (ghost buffer); thread
            // (lrow, lj) takes column lj of rows 1 + lrow, 1 + lrow + lrows, ...
            // of each strip.  The loads of kLoadBatch records are issued
            // before any of them is stored to LDS, so a strip costs about one
            // memory round trip rather than one per record.
            if (lrow < lrows) {
                constexpr int kLoadBatch = 4;
                // the next record: strip k, local row lr (none when the
                // thread's first row is past the strip: k = K)
                int k = 1 + lrow <= H + 1 ? 0 : K, lr = 1 + lrow;
                while (k < K) {
                    ulonglong2 rec[kLoadBatch];
                    float th[kLoadBatch];
                    int8_t sp[kLoadBatch];
                    int at[kLoadBatch];  // (strip << 16) | local row of record i, or -1
#pragma unroll
                    for (int i = 0; i < kLoadBatch; i++) {
                        at[i] = -1;
                        if (k < K) {
                            const size_t rep = (size_t)(r0 + k) * L * L + (size_t)(row0 + lr - 1) * L + lj;
                            const size_t gh = ((size_t)(r0 + k) * S + s) * L + lj;
                            const bool ghost = lr > H;
                            rec[i] = load_rec(ghost ? as_global<RngRec>(g.ghost_rng) + gh
                                                    : as_global<RngRec>(g.rng) + rep);
                            if (MODEL == MODEL_ISING)
                                sp[i] = ghost ? as_global<int8_t>(g.ghost_state)[gh] : as_global<int8_t>(g.state)[rep];
                            else
                                th[i] = ghost ? as_global<float>(g.ghost_state)[gh] : as_global<float>(g.state)[rep];
                            at[i] = (k << 16) | lr;
                            lr += lrows;
                            if (lr > H + 1) {
                                lr = 1 + lrow;
                                k++;
                            }
                        }
                    }
#pragma unroll
                    for (int i = 0; i < kLoadBatch; i++) {
                        if (at[i] < 0) continue;
                        const StripView v = view_of(at[i] >> 16);
                        const int off = ((at[i] & 0xffff) - 1) * L;
                        v.rng[off + ljc] = rec[i];
                        if (MODEL == MODEL_ISING) {
                            v.sp[off + L + ljc] = sp[i];
                        } else {
                            float sv, cv;
                            fast_sincosf(th[i], sv, cv);
                            v.cs[off + L + ljc] = make_float2(cv, sv);
                            v.th[off + ljc] = th[i];
                        }
                    }
                    // every load of the batch has been used: said explicitly, so
                    // that the next batch's loads are not each held back by a
                    // wait for the registers they reuse
                    __builtin_amdgcn_s_waitcnt(0);
                }
            }
            // ---- rows 0 and H+2: the prefetched words (polled again only if
            // a neighbour was late), or rows H and 2 of this strip itself
            // when it is a whole replica (S == 1)
            {
                auto put = [&](int qr, uint32_t payload) {
                    const StripView v = view_of(qr >> 1);
                    const int row = (qr & 1) ? H + 2 : 0;
                    if (MODEL == MODEL_ISING) {
                        v.sp[lds_index(row, lj, L)] = (int8_t)(uint8_t)payload;
                    } else {
                        float sv, cv;
                        fast_sincosf(__uint_as_float(payload), sv, cv);
                        v.cs[lds_index(row, lj, L)] = make_float2(cv, sv);
                    }
                };
                int bad = 0;
                const unsigned tag = (unsigned)n + 1u;
                if (h0) {
                    uint32_t payload = (uint32_t)hw0;
                    if ((unsigned)(hw0 >> 32) != tag && !poll_word(halo_src(hq0), tag, payload, a.error)) bad = 1;
                    put(hq0, payload);
                }
                if (h1) {
                    uint32_t payload = (uint32_t)hw1;
                    if ((unsigned)(hw1 >> 32) != tag && !poll_word(halo_src(hq1), tag, payload, a.error)) bad = 1;
                    put(hq1, payload);
                }
                if (bad) s_err = 1;
                __syncthreads();
                if (S == 1) {
                    if (lrow < lrows)
                        for (int qr = lrow; qr < 2 * K; qr += lrows) {
                            const StripView v = view_of(qr >> 1);
                            const int dst = (qr & 1) ? H + 2 : 0, src = (qr & 1) ? 2 : H;
                            if (MODEL == MODEL_ISING) {
                                v.sp[lds_index(dst, lj, L)] = v.sp[lds_index(src, lj, L)];
                            } else {
                                v.cs[lds_index(dst, lj, L)] = v.cs[lds_index(src, lj, L)];
                            }
                        }
                    __syncthreads();
                }
                if (s_err) {
                    if (tid == 0) atomicOr(a.error, BATCH_ERR_WATCHDOG);
                    failed = true;
                    break;
                }
            }

            // ---- colours 0, 1, 2, 3 in the reference order; colours 0/1
            // include the ghost row (colour row HH), whose moves are its
            // owner's and are not counted
            const StripView pv = view_of(pass_on ? pk : 0);
            const float pbeta = pass_on ? pt_beta[r0 + pk] : 0.0f;
            const float psig = MODEL == MODEL_XY_GAUSS && pass_on ? sig[pk] : 0.0f;
            const float *pthr = thr + 5 * (pass_on ? pk : 0);
            int n_acc = 0;
            for (int c = 0; c < 4; c++) {
                if (pass_on) {
                    const int nrows = c < 2 ? HH + 1 : HH;
                    for (int ri = prow; ri < nrows; ri += rpp) {
                        const int counted = !((c < 2) & (ri == HH));
                        if (site_update<MODEL>(pv, c, ri, tcol, L, pbeta, psig, J, delta, pthr, q_tab))
                            n_acc += counted;
                    }
                }
                __syncthreads();
            }
            if (pass_on && n_acc) atomicAdd(&moves[r0 + pk], (unsigned)n_acc);

            // ---- publish rows 2 and H (tag n + 2, slot (n + 1) & 1); write back
            if (lrow < lrows) {
                if (S > 1)
                    for (int qr = lrow; qr < 2 * K; qr += lrows) {
                        const int k = qr >> 1, which = qr & 1;
                        const StripView v = view_of(k);
                        const int lr = which ? H : 2;
                        const uint32_t payload = MODEL == MODEL_ISING
                                                     ? (uint32_t)(uint8_t)v.sp[lds_index(lr, lj, L)]
                                                     : __float_as_uint(v.th[lds_index(lr - 1, lj, L)]);
                        publish_word(halo_ptr(r0 + k, s, which, (n + 1) & 1) + lj, (unsigned)n + 2u, payload);
                    }
                for (int k = 0; k < K; k++) {
                    const StripView v = view_of(k);
                    const size_t rep = (size_t)(r0 + k) * L * L + (size_t)row0 * L + lj;
                    const auto rng_dst = as_global<RngRec>(g.rng) + rep;
                    const size_t gh = ((size_t)(r0 + k) * S + s) * L + lj;
#pragma unroll 2
                    for (int lr = 1 + lrow; lr <= H + 1; lr += lrows) {
                        const int off = (lr - 1) * L;
                        const ulonglong2 rec = v.rng[off + ljc];
                        if (lr > H) {
                            store_rec(as_global<RngRec>(g.ghost_rng) + gh, rec);
                            if (MODEL == MODEL_ISING)
                                as_global<int8_t>(g.ghost_state)[gh] = v.sp[off + L + ljc];
                            else
                                as_global<float>(g.ghost_state)[gh] = v.th[off + ljc];
                        } else {
                            store_rec(rng_dst + off, rec);
                            if (MODEL == MODEL_ISING)
                                as_global<int8_t>(g.state)[rep + off] = v.sp[off + L + ljc];
                            else
                                as_global<float>(g.state)[rep + off] = v.th[off + ljc];
                        }
                    }
                }
            }

            // ---- strip observables per replica (segmented reduction over the
            // replica's NT / K threads)
            if (do_exchange || do_sample) {
                long long ef = 0;
                double m[4] = {0.0, 0.0, 0.0, 0.0};
                const StripView v = view_of(ok);
                const int prl = TPR >= L ? TPR / L : 1;
                const int c0 = TPR >= L ? ot % L : ot, rr0 = TPR >= L ? ot / L : 0;
                const int cstep = TPR >= L ? L : TPR;
                if (rr0 < prl)
                    for (int lr = 1 + rr0; lr <= H; lr += prl)
                        for (int j = c0; j < L; j += cstep) site_observables<MODEL>(v, lr, j, L, do_sample, ef, m);
                for (int off = W >> 1; off > 0; off >>= 1) {
                    ef += __shfl_down(ef, off, W);
#pragma unroll
                    for (int c = 0; c < 4; c++) m[c] += __shfl_down(m[c], off, W);
                }
                if (ot % W == 0) {
                    const int slot = tid / W;
                    red_e[slot] = ef;
#pragma unroll
                    for (int c = 0; c < 4; c++) red[c][slot] = m[c];
                }
                __syncthreads();
                if (tid < K) {
                    const int per = TPR / W, first = tid * per;
                    long long E = red_e[first];
                    double M[4] = {red[0][first], red[1][first], red[2][first], red[3][first]};
                    for (int x = 1; x < per; x++) {
                        E += red_e[first + x];
#pragma unroll
                        for (int c = 0; c < 4; c++) M[c] += red[c][first + x];
                    }
                    const int r = r0 + tid;
                    if (do_exchange) {
                        const long long xi = sweep_global / g.exchange_interval;
                        uint64_t *dst = g.energy_words + ((((size_t)(xi & 1) * R + r) * S + s) * 2);
                        publish_word(dst, (unsigned)xi + 1u, (uint32_t)(unsigned long long)E);
                        publish_word(dst + 1, (unsigned)xi + 1u, (uint32_t)((unsigned long long)E >> 32));
                    }
                    if (do_sample && sidx < a.max_samples) {
                        double *dst = a.sample_partials[w.group] + (((size_t)sidx * R + r) * S + s) * kChannels;
                        dst[0] = energy_from_fixed(MODEL, J, E);
#pragma unroll
                        for (int c = 0; c < 4; c++) dst[c + 1] = M[c];
                    }
                }
            }
            __syncthreads();
        }
        if (failed) break;

        if (do_exchange) {
            flush_moves();  // moves so far belong to the current temperatures
            for (int k = tid; k < R; k += NT) psum[k] = 0;
            __syncthreads();
            // every strip energy of the group's replicas, added as integers
            const long long xi = sweep_global / g.exchange_interval;
            const unsigned tag = (unsigned)xi + 1u;
            const uint64_t *words = g.energy_words + (size_t)(xi & 1) * R * S * 2;
            int bad = 0;
            for (int k = tid; k < R * S; k += NT) {
                uint32_t lo, hi;
                if (!poll_word(words + 2 * (size_t)k, tag, lo, a.error) ||
                    !poll_word(words + 2 * (size_t)k + 1, tag, hi, a.error)) {
                    bad = 1;
                    break;
                }
                atomicAdd(reinterpret_cast<unsigned long long *>(psum + k / S),
                          ((unsigned long long)hi << 32) | lo);
            }
            if (bad) s_err = 1;
            __syncthreads();
            if (s_err) {
                if (tid == 0) atomicOr(a.error, BATCH_ERR_WATCHDOG);
                failed = true;
                break;
            }
            for (int k = tid; k < R; k += NT) pt_e[k] = energy_from_fixed(MODEL, J, psum[k]);
            __syncthreads();
            const int P = R - 1;
            for (int left = parity + 2 * tid; left < R - 1; left += 2 * NT) {
                const int right = left + 1;
                const int wa = pt_t2w[left], wb = pt_t2w[right];
                const double ea = pt_e[wa], eb = pt_e[wb];
                const bool finite = isfinite(ea) && isfinite(eb);
                bool accept = false;
                if (finite) {
                    const double lr = ((double)pt_beta[wa] - (double)pt_beta[wb]) * (ea - eb);
                    Rng gg;
                    gg.s0 = pt_g[2 * left];
                    gg.s1 = pt_g[2 * left + 1];
                    const double u = gg.uniform_f64();
                    pt_g[2 * left] = gg.s0;
                    pt_g[2 * left + 1] = gg.s1;
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
                if (left % g.wg_count == lw) {
                    pt_cnt[left] += 1;
                    if (!finite) pt_cnt[2 * P + left] += 1;
                    if (accept) pt_cnt[P + left] += 1;
                }
            }
            __syncthreads();
            for (int wk = lw + tid * g.wg_count; wk < R; wk += NT * g.wg_count) {
                const int tt = pt_w2t[wk];
                if (tt != 0 && tt != R - 1) continue;
                const long long origin = pt_trips[wk];
                if (origin < 0) {
                    pt_trips[wk] = tt;
                    pt_trips[2 * R + wk] = sweep_global;
                } else if (tt != origin) {
                    pt_trips[R + wk] = 1;
                } else if (pt_trips[R + wk]) {
                    pt_trips[3 * R + wk] += 1;
                    pt_trips[4 * R + wk] += sweep_global - pt_trips[2 * R + wk];
                    pt_trips[2 * R + wk] = sweep_global;
                    pt_trips[R + wk] = 0;
                }
            }
            if (R > 2) parity = 1 - parity;
            exchanged = true;
            __syncthreads();
        }
        if (do_sample && lw == 0 && sidx < a.max_samples)
            for (int k = tid; k < R; k += NT) a.sample_w2t[w.group][(size_t)sidx * R + k] = pt_w2t[k];
    }

    flush_moves();
    if (lw == 0 && use_pt) {
        for (int k = tid; k < R; k += NT) {
            gc.betas[k] = pt_beta[k];
            gc.t2w[k] = pt_t2w[k];
            gc.w2t[k] = pt_w2t[k];
            if (exchanged) gc.energies[k] = pt_e[k];
        }
        for (int k = tid; k < 2 * (R - 1); k += NT) gc.pt_rng[k] = pt_g[k];
        if (tid == 0) *gc.parity = parity;
    }
    if (use_pt) {  // the kept walkers' records and pairs' counters
        const int P = R - 1;
        for (int wk = lw + tid * g.wg_count; wk < R; wk += NT * g.wg_count)
            for (int row = 0; row < 5; row++) gc.trips[row * R + wk] = pt_trips[row * R + wk];
        for (int left = lw + tid * g.wg_count; left < P; left += NT * g.wg_count)
            for (int row = 0; row < 3; row++) gc.counts[row * P + left] += pt_cnt[row * P + left];
    }
    if (blockIdx.x == 0 && tid == 0) a.clock[0] = sweep0 + a.nsweeps;
}

template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) void pack_kernel(PackArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    __shared__ long long red_e[NT / 16];
    __shared__ double red[4][NT / 16];
    __shared__ float4 q_tab[kNormalCells];  // Gaussian XY (ordered by pack_run's first barrier)
    __shared__ int s_err;
    const PackWork w = a.work[blockIdx.x];
    const PackGroup *g = a.groups + w.group;
    const int model = g->model;
    if (model == MODEL_ISING) {
        pack_run<MODEL_ISING, NT>(a, g, w, smem, red_e, red, q_tab, s_err);
    } else if (model == MODEL_XY_FAST) {
        pack_run<MODEL_XY_FAST, NT>(a, g, w, smem, red_e, red, q_tab, s_err);
    } else {
        normal_table_to_lds(q_tab);
        pack_run<MODEL_XY_GAUSS, NT>(a, g, w, smem, red_e, red, q_tab, s_err);
    }
}

// Before each launch: ghost rows from the lattice, halo words of sweep 0
// (tag 1, slot 0; slot 1 cleared), exchange words cleared, the launch's base
// sweep, the error word.  blockIdx.y = group.
__global__ void pack_prologue_kernel(PackArgs a) {
    const PackGroup g = a.groups[blockIdx.y];
    const int L = g.L, S = g.S, H = g.H, R = g.R;
    const size_t n = (size_t)R * S * L;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) {
        const int r = (int)(k / ((size_t)S * L));
        const int rem = (int)(k - (size_t)r * S * L);
        const int s = rem / L, j = rem - (rem / L) * L;
        const size_t rep = (size_t)r * L * L;
        const size_t gk = rep + (size_t)(((s + 1) % S) * H) * L + j;
        reinterpret_cast<ulonglong2 *>(g.ghost_rng)[k] = reinterpret_cast<const ulonglong2 *>(g.rng)[gk];
        const size_t k2 = rep + (size_t)(s * H + 1) * L + j, kH = rep + (size_t)(s * H + H - 1) * L + j;
        uint32_t p2, pH;
        if (g.model == MODEL_ISING) {
            const int8_t *st = reinterpret_cast<const int8_t *>(g.state);
            reinterpret_cast<int8_t *>(g.ghost_state)[k] = st[gk];
            p2 = (uint32_t)(uint8_t)st[k2];
            pH = (uint32_t)(uint8_t)st[kH];
        } else {
            const float *st = reinterpret_cast<const float *>(g.state);
            reinterpret_cast<float *>(g.ghost_state)[k] = st[gk];
            p2 = __float_as_uint(st[k2]);
            pH = __float_as_uint(st[kH]);
        }
        uint64_t *h = g.halo + (((size_t)r * S + s) * 4) * L + j;  // [which][slot][L]
        h[0] = ((uint64_t)1 << 32) | p2;  // row 2, slot 0
        h[L] = 0;                          // row 2, slot 1
        h[2 * L] = ((uint64_t)1 << 32) | pH;
        h[3 * L] = 0;
    }
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < (size_t)4 * R * S; k += stride)
        g.energy_words[k] = 0;
    if (blockIdx.x == 0) {
        if (blockIdx.y == 0 && threadIdx.x == 0) {
            *a.error = 0;
            a.clock[1] = a.clock[0];
        }
    }
}

// ------------------------------------------------------------------ host side

struct GroupHost {
    PackGroup g;
    void *halo, *ghost_rng, *ghost_state, *acc;
};

struct Pack {
    int threads = 256;
    std::vector<GroupHost> groups;
    std::vector<PackWork> work;
    PackGroup *d_groups = nullptr;
    PackWork *d_work = nullptr;
    long long *d_clock = nullptr;
    int *d_error = nullptr;
    size_t lds = 0;
    bool dirty = true;  // the device group table needs an upload
};

static const void *pack_kernel_ptr(int threads) {
    return threads == 512 ? reinterpret_cast<const void *>(&pack_kernel<512>)
                          : reinterpret_cast<const void *>(&pack_kernel<256>);
}

static int pack_check_group(int threads, int model, int L, int R, int H, int K) {
    MC_REQUIRE(model == MODEL_ISING || model == MODEL_XY_FAST || model == MODEL_XY_GAUSS,
               "unknown batch model %d", model);
    MC_REQUIRE(L >= 4 && L % 2 == 0 && L <= threads, "packed L must be even, >= 4 and <= %d, got %d", threads,
               L);
    MC_REQUIRE(H >= 4 && H % 2 == 0 && L % H == 0, "strip height %d must be even, >= 4 and divide L=%d", H, L);
    MC_REQUIRE(K >= 1 && K <= threads / 16 && (K & (K - 1)) == 0 && R % K == 0,
               "replicas per task K=%d must be a power of two <= %d dividing R=%d", K, threads / 16, R);
    MC_REQUIRE((threads / (L / 2)) / K >= 1, "K=%d strips of L=%d do not fit a %d-thread pass", K, L, threads);
    MC_REQUIRE(K * L <= threads, "K=%d replicas of L=%d exceed one %d-thread row map", K, L, threads);
    MC_REQUIRE((long long)L * L <= (1LL << 20), "at most 2^20 sites per replica");
    return MC_OK;
}

static void pack_free(Pack *p) {
    for (auto &gh : p->groups) {
        (void)hipFree(gh.halo);
        (void)hipFree(gh.ghost_rng);
        (void)hipFree(gh.ghost_state);
        (void)hipFree(gh.acc);
    }
    p->groups.clear();
    (void)hipFree(p->d_groups);
    (void)hipFree(p->d_work);
    (void)hipFree(p->d_clock);
    (void)hipFree(p->d_error);
}

static int pack_upload(Pack *p) {
    if (!p->dirty) return MC_OK;
    std::vector<PackGroup> tab;
    for (auto &gh : p->groups) tab.push_back(gh.g);
    MC_TRY_HIP(hipMemcpy(p->d_groups, tab.data(), tab.size() * sizeof(PackGroup), hipMemcpyHostToDevice));
    p->dirty = false;
    return MC_OK;
}

static int pack_launch(Pack *p, hipStream_t stream, int nsweeps, int sample_interval, long long sample_first,
                       int max_samples, double *const *sample_partials, int32_t *const *sample_w2t) {
    PackArgs a;
    for (int i = 0; i < kMaxPackGroups; i++) {
        const bool have = max_samples > 0 && i < (int)p->groups.size();
        a.sample_partials[i] = have ? sample_partials[i] : nullptr;
        a.sample_w2t[i] = have ? sample_w2t[i] : nullptr;
    }
    a.groups = p->d_groups;
    a.work = p->d_work;
    a.ngroups = (int)p->groups.size();
    a.nsweeps = nsweeps;
    a.clock = p->d_clock;
    a.sample_interval = sample_interval;
    a.sample_first = sample_first;
    a.max_samples = max_samples;
    a.error = p->d_error;
    size_t most = 0;
    for (auto &gh : p->groups) most = std::max(most, (size_t)gh.g.R * gh.g.S * gh.g.L);
    const unsigned bx = (unsigned)std::min<size_t>((most + 255) / 256, 256);
    hipLaunchKernelGGL(pack_prologue_kernel, dim3(bx, (unsigned)p->groups.size()), dim3(256), 0, stream, a);
    MC_TRY_HIP(hipGetLastError());
    if (p->threads == 512)
        hipLaunchKernelGGL(pack_kernel<512>, dim3((unsigned)p->work.size()), dim3(512), p->lds, stream, a);
    else
        hipLaunchKernelGGL(pack_kernel<256>, dim3((unsigned)p->work.size()), dim3(256), p->lds, stream, a);
    MC_TRY_HIP(hipGetLastError());
    return MC_OK;
}

static int pack_prepare(Pack *p, int nsweeps, int sample_interval, int max_samples, double *const *sample_partials,
                        int32_t *const *sample_w2t) {
    MC_REQUIRE(!p->groups.empty() && !p->work.empty(), "the pack has no groups or no work table");
    MC_REQUIRE(nsweeps >= 0 && sample_interval >= 0 && max_samples >= 0, "invalid sweep or sampling arguments");
    int dev = 0, cus = 0, per_cu = 0;
    MC_TRY_HIP(hipGetDevice(&dev));
    MC_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (p->lds > 64 * 1024)
        MC_TRY_HIP(hipFuncSetAttribute(pack_kernel_ptr(p->threads), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)p->lds));
    MC_TRY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pack_kernel_ptr(p->threads), p->threads,
                                                            p->lds));
    MC_REQUIRE((long long)p->work.size() <= (long long)per_cu * cus,
               "%zu workgroups cannot be co-resident (%d per CU x %d CUs at %zu bytes of LDS)", p->work.size(),
               per_cu, cus, p->lds);
    for (size_t i = 0; max_samples > 0 && i < p->groups.size(); i++)
        MC_REQUIRE(sample_partials && sample_w2t && sample_partials[i] && sample_w2t[i],
                   "group %zu needs sample buffers", i);
    return pack_upload(p);
}

}  // namespace mc

using namespace mc;

extern "C" {

int mc_pack_lds_bytes(int model, int L, int R, int H, int K, size_t *bytes) {
    int rc = pack_check_group(512, model, L, R, H, K);
    if (rc) return rc;
    *bytes = pack_lds(model, L, R, H, K).total;
    return MC_OK;
}

int mc_pack_occupancy(int threads, size_t lds_bytes, int *max_workgroups) {
    MC_REQUIRE(threads == 256 || threads == 512, "workgroup size must be 256 or 512, got %d", threads);
    hipFuncAttributes attr;
    MC_TRY_HIP(hipFuncGetAttributes(&attr, pack_kernel_ptr(threads)));
    if (lds_bytes + attr.sharedSizeBytes > 160 * 1024) {  // dynamic plus static LDS of a CU
        *max_workgroups = 0;
        return MC_OK;
    }
    if (lds_bytes > 64 * 1024)
        MC_TRY_HIP(hipFuncSetAttribute(pack_kernel_ptr(threads), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds_bytes));
    int per_cu = 0, dev = 0, cus = 0;
    MC_TRY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pack_kernel_ptr(threads), threads, lds_bytes));
    MC_TRY_HIP(hipGetDevice(&dev));
    MC_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    *max_workgroups = per_cu * cus;
    return MC_OK;
}

int mc_pack_create(int threads, void **pack) {
    MC_REQUIRE(threads == 256 || threads == 512, "workgroup size must be 256 or 512, got %d", threads);
    Pack *p = new Pack();
    p->threads = threads;
    if (hipMalloc(&p->d_clock, 2 * sizeof(long long)) != hipSuccess ||
        hipMalloc(&p->d_error, sizeof(int)) != hipSuccess ||
        hipMemset(p->d_clock, 0, 2 * sizeof(long long)) != hipSuccess ||
        hipMemset(p->d_error, 0, sizeof(int)) != hipSuccess) {
        pack_free(p);
        delete p;
        mc::set_error(MC_ERR_HIP, "mc_pack_create: device allocation failed");
        return MC_ERR_HIP;
    }
    *pack = p;
    return MC_OK;
}

int mc_pack_destroy(void *pack) {
    if (!pack) return MC_OK;
    Pack *p = static_cast<Pack *>(pack);
    (void)hipDeviceSynchronize();
    pack_free(p);
    delete p;
    return MC_OK;
}

int mc_pack_add_group(void *pack, int model, int L, int R, int H, int K, int exchange_interval, float J,
                      float delta, void *state, uint64_t *rng, float *betas, const float *sigmas,
                      unsigned long long *accepted, int *parity, double *energies, int32_t *w2t, int32_t *t2w,
                      uint64_t *pt_rng, long long *counts, long long *trips, int *group_index) {
    Pack *p = static_cast<Pack *>(pack);
    MC_REQUIRE(p, "null pack");
    MC_REQUIRE(p->work.empty(), "groups must be added before the work table is set");
    MC_REQUIRE((int)p->groups.size() < kMaxPackGroups, "a pack holds at most %d groups", kMaxPackGroups);
    int rc = pack_check_group(p->threads, model, L, R, H, K);
    if (rc) return rc;
    MC_REQUIRE(exchange_interval >= 0, "exchange_interval must be nonnegative");
    MC_REQUIRE(exchange_interval == 0 || R >= 2, "exchanges need at least two replicas");
    MC_REQUIRE(model != MODEL_XY_GAUSS || sigmas, "the Gaussian XY model needs per-temperature sigmas");
    MC_REQUIRE(state && rng && betas, "state, rng and betas are required");
    MC_REQUIRE(exchange_interval == 0 || (parity && energies && w2t && t2w && pt_rng && counts && trips),
               "exchanges need the PT arrays");
    GroupHost gh{};
    PackGroup &g = gh.g;
    g.model = model; g.L = L; g.R = R; g.H = H; g.S = L / H; g.K = K;
    g.ntasks = (R / K) * g.S;
    g.exchange_interval = exchange_interval; g.J = J; g.delta = delta;
    g.state = state; g.rng = rng; g.betas = betas; g.sigmas = sigmas; g.accepted = accepted;
    g.parity = parity; g.energies = energies; g.w2t = w2t; g.t2w = t2w; g.pt_rng = pt_rng;
    g.counts = counts; g.trips = trips;
    const size_t rows = (size_t)R * g.S * L;
    const size_t sb = model == MODEL_ISING ? 1 : 4;
    if (hipMalloc(&gh.halo, rows * 4 * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc(&gh.ghost_rng, rows * 16) != hipSuccess || hipMalloc(&gh.ghost_state, rows * sb) != hipSuccess ||
        hipMalloc(&gh.acc, (size_t)4 * R * (L / H) * sizeof(uint64_t)) != hipSuccess) {
        (void)hipFree(gh.halo);
        (void)hipFree(gh.ghost_rng);
        (void)hipFree(gh.ghost_state);
        (void)hipFree(gh.acc);
        mc::set_error(MC_ERR_HIP, "mc_pack_add_group: device allocation failed");
        return MC_ERR_HIP;
    }
    g.halo = static_cast<uint64_t *>(gh.halo);
    g.ghost_rng = static_cast<uint64_t *>(gh.ghost_rng);
    g.ghost_state = gh.ghost_state;
    g.energy_words = static_cast<uint64_t *>(gh.acc);
    *group_index = (int)p->groups.size();
    p->groups.push_back(gh);
    p->lds = std::max(p->lds, (size_t)pack_lds(model, L, R, H, K).total);
    p->dirty = true;
    return MC_OK;
}

// work: nworkgroups rows of (group, first task, task count).  Every task of
// every group must appear exactly once, and a group's workgroups must be
// consecutive.
int mc_pack_set_work(void *pack, const int *work, int nworkgroups) {
    Pack *p = static_cast<Pack *>(pack);
    MC_REQUIRE(p && !p->groups.empty(), "add groups first");
    MC_REQUIRE(p->work.empty(), "the work table is already set");
    MC_REQUIRE(nworkgroups > 0, "need at least one workgroup");
    const int G = (int)p->groups.size();
    std::vector<int> first(G, -1), count(G, 0), next(G, 0);
    std::vector<PackWork> tab(nworkgroups);
    for (int i = 0; i < nworkgroups; i++) {
        const int gi = work[3 * i], t0 = work[3 * i + 1], n = work[3 * i + 2];
        MC_REQUIRE(gi >= 0 && gi < G, "work row %d names group %d", i, gi);
        MC_REQUIRE(n >= 1, "work row %d has no tasks", i);
        if (first[gi] < 0) first[gi] = i;
        MC_REQUIRE(first[gi] + count[gi] == i, "the workgroups of group %d are not consecutive", gi);
        MC_REQUIRE(t0 == next[gi], "work row %d: group %d's tasks must be listed in order", i, gi);
        next[gi] += n;
        count[gi]++;
        tab[i] = PackWork{gi, t0, n, 0};
    }
    for (int gi = 0; gi < G; gi++) {
        PackGroup &g = p->groups[gi].g;
        MC_REQUIRE(next[gi] == g.ntasks, "group %d: %d of %d tasks assigned", gi, next[gi], g.ntasks);
        g.wg_first = first[gi];
        g.wg_count = count[gi];
    }
    MC_TRY_HIP(hipMalloc(&p->d_work, tab.size() * sizeof(PackWork)));
    MC_TRY_HIP(hipMemcpy(p->d_work, tab.data(), tab.size() * sizeof(PackWork), hipMemcpyHostToDevice));
    MC_TRY_HIP(hipMalloc(&p->d_groups, p->groups.size() * sizeof(PackGroup)));
    p->work = tab;
    p->dirty = true;
    return MC_OK;
}

int mc_pack_set_sweeps(void *pack, long long sweeps) {
    Pack *p = static_cast<Pack *>(pack);
    MC_REQUIRE(p, "null pack");
    const long long c[2] = {sweeps, sweeps};
    MC_TRY_HIP(hipMemcpy(p->d_clock, c, sizeof(c), hipMemcpyHostToDevice));
    return MC_OK;
}

int mc_pack_sweeps(void *pack, long long *sweeps) {
    Pack *p = static_cast<Pack *>(pack);
    MC_REQUIRE(p, "null pack");
    MC_TRY_HIP(hipMemcpy(sweeps, p->d_clock, sizeof(long long), hipMemcpyDeviceToHost));
    return MC_OK;
}

int mc_pack_error(void *pack, int *error) {
    Pack *p = static_cast<Pack *>(pack);
    MC_REQUIRE(p, "null pack");
    MC_TRY_HIP(hipMemcpy(error, p->d_error, sizeof(int), hipMemcpyDeviceToHost));
    return MC_OK;
}

int mc_pack_run(void *pack, void *stream, int nsweeps, int sample_interval, long long sample_first, int max_samples,
                double *const *sample_partials, int32_t *const *sample_w2t) {
    Pack *p = static_cast<Pack *>(pack);
    MC_REQUIRE(p, "null pack");
    int rc = pack_prepare(p, nsweeps, sample_interval, max_samples, sample_partials, sample_w2t);
    if (rc) return rc;
    if (nsweeps == 0) return MC_OK;
    return pack_launch(p, as_stream(stream), nsweeps, sample_interval, sample_first, max_samples, sample_partials,
                       sample_w2t);
}

int mc_pack_capture(void *pack, void *stream, int nsweeps, int sample_interval, long long sample_first,
                    int max_samples, double *const *sample_partials, int32_t *const *sample_w2t, void **graph) {
    Pack *p = static_cast<Pack *>(pack);
    MC_REQUIRE(p, "null pack");
    MC_REQUIRE(stream, "graph capture needs a created stream");
    MC_REQUIRE(nsweeps > 0, "a captured launch needs sweeps");
    int rc = pack_prepare(p, nsweeps, sample_interval, max_samples, sample_partials, sample_w2t);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    MC_TRY_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = pack_launch(p, s, nsweeps, sample_interval, sample_first, max_samples, sample_partials, sample_w2t);
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(s, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    MC_TRY_HIP(end);
    hipGraphExec_t exec = nullptr;
    const hipError_t inst = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    MC_TRY_HIP(inst);
    *graph = exec;
    return MC_OK;
}

int mc_graph_launch(void *graph, void *stream) {
    MC_REQUIRE(graph, "null graph");
    MC_TRY_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(graph), as_stream(stream)));
    return MC_OK;
}

int mc_graph_destroy(void *graph) {
    if (graph) MC_TRY_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph)));
    return MC_OK;
}

}  // extern "C"
