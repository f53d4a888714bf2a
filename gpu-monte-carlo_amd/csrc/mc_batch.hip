// Persistent batched engine: a whole run (many sweeps, replica exchanges and
// samples) in one launch of co-resident workgroups.
//
// Decomposition: each replica lattice (L x L) is cut into S = L / H strips of
// H full rows; one workgroup (256 or 512 threads) owns one strip for the
// whole launch.
// Everything the strip needs lives in LDS, rows stored colour-split (even
// columns, then odd columns) so every access of a colour step is unit-stride:
//   local row 0      bottom row of the strip above (received)
//   rows 1..H        own rows: state, RNG streams, cos/sin (or spins)
//   row H+1          ghost: the top row of the strip below, with a copy of its
//                    RNG streams, updated here in lock-step with its owner
//   row H+2          second row of the strip below (received)
// Global memory is touched only at launch start/end, for the two received
// rows per sweep and for exchanges and samples.
//
// The ghost row removes every dependence on a neighbour's current sweep: the
// top row of strip s+1 depends on the bottom row of strip s (here) and on row
// 2 of strip s+1 from the previous sweep, and it is all that the bottom row
// of strip s needs from below.  Its update repeats the owner's operations on
// the same inputs and streams, so both copies agree bit for bit.  So a strip
// only waits for data of the previous sweep -- the bottom row of s-1 and row 2
// of s+1 -- published at the end of that sweep and used by the boundary sites
// of the third and fourth colour passes, after two passes of interior rows
// that hide the hand-off latency.  A strip can run
// at most one sweep ahead of its neighbours, so halo slots are double-buffered
// by sweep parity.  Each halo word carries its payload (angle or spin) and the
// sweep it belongs to, so the receiver polls the data itself: one L2 round
// trip per hand-off, no flags.
//
// Strips of one replica are placed on one XCD.  Every `exchange_interval`
// sweeps each strip publishes its integer energy partial as two tagged words
// (low and high halves) under its walker's ladder temperature.  A pair's
// decision depends only on the two walkers' energies, the two ladder betas and
// the pair's own PT stream, so each workgroup gathers just its own walker's
// and its partner's partials and decides that pair; the strips of both walkers
// reach the same decision, and nothing waits for the other pairs (no grid
// barrier).  Every workgroup advances every attempted pair's stream by one
// draw (energies are integer sums here, always finite), so the streams stay
// identical everywhere.  The energy words form a ring of four exchanges: a
// walker can be at most two exchanges past a partner that has not read its
// words yet, because the same two ladder slots pair up again two exchanges
// later.  Each walker's first strip keeps its round-trip record and counts
// the pairs it is the left member of; the PT streams go back from workgroup 0.
// Every workgroup reads the shared PT state (betas, maps, PT streams, parity)
// when it starts.  The state at the end of the launch goes to a staging area
// of the workspace, and a small kernel queued after the launch copies it in
// place, so no workgroup can read state of the launch's end, however late it
// starts.  (A start counter with an end-of-launch wait for it does the same
// inside the launch, but its code alone -- never executed in a launch without
// exchanges -- cost the sweep loop 5% on MI355X through register allocation.)
//
// Every wait is bounded: past a watchdog count a workgroup records an error
// and leaves, so a broken invariant ends the launch instead of hanging it.
//
// Models:
//  * MODEL_ISING: the reference's Ising update, bit-identical to the
//    per-colour kernels (thresholds exp(-beta dE) recomputed per beta change).
//  * MODEL_XY_FAST: the XY update with dE from cached (cos, sin) of every site
//    and the bit-reproducible float32 sincos/exp of mc_fastmath.h.
//  * MODEL_XY_GAUSS: the same with a Gaussian proposal theta + sigma_t z,
//    sigma_t per ladder temperature (it follows the walker's temperature
//    through exchanges); z and the acceptance uniform come from one 64-bit
//    draw per site (gauss_draw_k in mc_fastmath.h).
// Accepted moves are counted per ladder temperature when `accepted` is given.
#include <algorithm>
#include <mutex>
#include <new>
#include <vector>

#include "mc_strip.h"

namespace mc {

struct BatchArgs {
    int model, L, R, H, S;
    void *state;                // int8 or float32 (R, L, L)
    uint64_t *rng;              // R*L*L records
    float *betas;               // (R) walker betas, updated at exchanges
    float J, delta;
    const float *sigmas;        // (R) per temperature, MODEL_XY_GAUSS
    unsigned long long *accepted;  // (R) per temperature, may be NULL
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
    // this launch's first exchange and sample sweeps (n; nsweeps if none) and
    // their indices, so the sweep loop needs no 64-bit division
    long long x_next, x_index, s_next, s_index;
    int max_samples;
    double *sample_partials;    // [max_samples][R][S][kChannels]
    int32_t *sample_w2t;        // [max_samples][R]
    // synchronisation
    uint64_t *halo;             // [R][S][2 (row 2, row H)][2 (sweep parity)][L] words
                                //   (tag << 32 | payload)
    uint64_t *energy;           // [4 (exchange mod 4)][R (ladder temperature)][S][2 (lo, hi)] words
                                //   (tag << 32 | half of the int64 energy partial)
    int *error;                 // the watchdog code
    // Words written at the end of sweep n of this launch carry the tag
    // tag0 + n + 1.  tag0 is at least every tag an earlier launch wrote into
    // this workspace since it was zeroed, so a word left over from an earlier
    // launch never matches a tag this launch waits for, and the workspace
    // needs no clearing between launches.
    unsigned tag0;
    unsigned recv_tag0;         // tag0, or tags never published (test hook)
    unsigned watchdog;          // polls before a wait gives up
};

// The staging area of the end-of-launch PT state, after the energy words of
// the workspace: PT streams (2 (R - 1) records), then walker betas, walker
// to temperature and temperature to walker maps (R each) and the parity.
struct BatchStage {
    uint64_t *pt_rng;
    float *betas;
    int32_t *w2t, *t2w;
    int *parity;
};
__host__ __device__ inline size_t batch_stage_words(int R) {
    return (size_t)2 * (R > 1 ? R - 1 : 0) + ((size_t)3 * R + 1 + 1) / 2;
}
__host__ __device__ inline BatchStage batch_stage(uint64_t *energy, int R, int S) {
    BatchStage st;
    st.pt_rng = energy + (size_t)8 * R * S;
    st.betas = reinterpret_cast<float *>(st.pt_rng + 2 * (R > 1 ? R - 1 : 0));
    st.w2t = reinterpret_cast<int32_t *>(st.betas + R);
    st.t2w = st.w2t + R;
    st.parity = st.t2w + R;
    return st;
}

// Item q of pass p (the five-pass schedule below): colour c and colour row
// ri; false if the pass has fewer items.
__device__ __forceinline__ bool pass_item(const int p, const int q, const int HH, int &c, int &ri) {
    switch (p) {
    case 0:  // c0 of colour rows 1..HH-1 (local rows 3..H-1)
    case 1:  // c1 of the same rows
        c = p;
        ri = 1 + q;
        return q < HH - 1;
    case 2:  // c0 of colour rows 0 and HH (ghost); c2 of colour rows 1..HH-2
        c = q < 2 ? 0 : 2;
        ri = q == 0 ? 0 : (q == 1 ? HH : q - 1);
        return q < HH;
    case 3:  // c1 of colour rows 0 and HH; c2 of colour rows 0 and HH-1
        c = q < 2 ? 1 : 2;
        ri = (q == 0 || q == 2) ? 0 : (q == 1 ? HH : HH - 1);
        return q < 4;
    default:  // c3 of colour rows 0..HH-1
        c = 3;
        ri = q;
        return q < HH;
    }
}

// Sum of S strip partials stored as (lo, hi) halves
__device__ __forceinline__ long long strip_total(const uint32_t *w, int S) {
    long long t = 0;
    for (int q = 0; q < S; q++) t += (long long)(((unsigned long long)w[2 * q + 1] << 32) | w[2 * q]);
    return t;
}

template <int MODEL, int NT, int MAXP>
__global__ __launch_bounds__(NT) void batch_kernel(BatchArgs a) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    const int L = a.L, H = a.H, S = a.S, R = a.R;
    const int half = L >> 1, HH = H >> 1;
    const int tid = threadIdx.x;
    const unsigned nwg = gridDim.x;

    // XCD-aware placement: all strips of a replica on one XCD when possible
    int lin = blockIdx.x;
    if (nwg % 8 == 0 && (nwg / 8) % S == 0) lin = (blockIdx.x % 8) * (nwg / 8) + blockIdx.x / 8;
    const int r = lin / S, s = lin % S;
    const int s_up = (s + S - 1) % S, s_dn = (s + 1) % S;

    // ---- LDS carve-up (the RNG streams live in registers)
    const StripView view = strip_view(smem, L, H, false);
    float2 *lcs = view.cs;  // rows 0..H+2: (cos, sin)
    float *lth = view.th;   // rows 1..H+1
    int8_t *lsp = view.sp;  // rows 0..H+2 (Ising)
    unsigned char *pt_base = smem + ((strip_bytes(MODEL, L, H, false) + 15) & ~(size_t)15);
    uint64_t *pt_g = reinterpret_cast<uint64_t *>(pt_base);        // 2 (R-1): PT streams
    float *ladder = reinterpret_cast<float *>(pt_g + 2 * (R - 1));  // R: beta of each ladder temperature
    float *sig_tab = ladder + R;                                    // R: proposal width of each temperature
    __shared__ long long my_trips[5];  // this walker's round-trip record (strip 0)
    __shared__ int x_acc;              // the exchange decision of this walker's pair
    __shared__ double x_e;             // this walker's energy at the exchange
    __shared__ int s_err;              // a wait of this workgroup expired
    __shared__ double red[kChannels][NT / 64];
    __shared__ long long red_e[NT / 64];
    __shared__ float thr[5];
    __shared__ int red_i[NT / 64];
    __shared__ float4 q_tab[MODEL == MODEL_XY_GAUSS ? kNormalCells : 1];  // normal quantile table of gauss_draw_k

    const size_t rep = (size_t)r * L * L;
    const int row0 = s * H;  // first global row of the strip
    const bool use_pt = a.exchange_interval > 0;

    // ---- every load that no other load feeds is issued before the strip's:
    // the walker's PT state, the first chunk of the shared PT streams, ladder
    // and widths (into registers, stored to LDS after the strip), and the RNG
    // streams of the thread's sites.  Loads complete in order, so the strip
    // load's wait covers them: one memory round trip at the start instead of
    // one per dependent group.
    float my_beta = a.betas[r];
    int my_t = use_pt ? a.w2t[r] : r;
    int parity = use_pt ? *a.parity : 0;
    float pt_beta = 0.0f, pt_sigma = 0.0f;
    int pt_walker_t = 0;
    uint64_t pt_stream = 0;
    long long pt_trip = 0;
    if (use_pt && tid < R) {  // walker tid's beta and ladder temperature (betas are in walker order)
        pt_beta = a.betas[tid];
        pt_walker_t = a.w2t[tid];
    }
    if (use_pt && tid < 2 * (R - 1)) pt_stream = a.pt_rng[tid];
    if (use_pt && s == 0 && tid < 5) pt_trip = a.trips[tid * R + r];
    if (MODEL == MODEL_XY_GAUSS && tid < R) pt_sigma = a.sigmas[tid];

    const float J = a.J, delta = a.delta;
    int n_acc = 0;  // accepted moves of this thread since the last flush
    // fixed thread -> (column, colour-row offset) map of every pass: a thread
    // takes items trow, trow + rows_per_pass, ... (at most MAXP) of each pass
    const int tcol = tid % half, trow = tid / half, rows_per_pass = NT / half;
    // the observables' sites tid, tid + NT, ... of rows 1..H, stepped by
    // (NT / L, NT % L) rows and columns instead of a division per site
    const int obs_i = tid / L, obs_j = tid % L, obs_di = NT / L, obs_dj = NT % L;

    // ---- the thread's sites: LDS positions and RNG streams in registers for
    // the whole launch (the ghost row's streams are a copy of its owner's)
    constexpr int NP = 5;
    // The halo words of sweep n are loaded once, in the draw phase (after the
    // draws of the first kPrefetchPass passes' sites), and consumed after two
    // colour passes; only a word that had not arrived yet is polled again.
    // Measured on MI355X (XY L=128 R=32 Gaussian, H=16, 512 threads), the
    // position: before any draws 160.8, after 1, 2, 3, 4 passes' draws 167.3,
    // 166.0, 169.6, 166.4, after all draws 166.0 site-updates/ns.
    constexpr int kPrefetchPass = 3;
    int slot_self[NP][MAXP], slot_lf[NP][MAXP], slot_rt[NP][MAXP];
    bool slot_on[NP][MAXP], slot_own[NP][MAXP];
    // Colours 0 and 1 of local row 1 read row 0 (the strip above), and of
    // the ghost row H+1 read row H+2 (the strip below); these boundary sites
    // are items 0 and 1 of passes 2 and 3.  Each takes its one received word
    // itself (halo word offset slot_hw, without the parity term), loaded
    // ahead into slot_pre: the received rows are read by no other site, so
    // they never go through LDS and need no barrier of their own.
    int slot_bnd[2][MAXP];     // 0: none, 1: up from the strip above, 2: down from the strip below
    int slot_hw[2][MAXP];
    uint64_t slot_pre[2][MAXP];
    uint64_t slot_s0[NP][MAXP], slot_s1[NP][MAXP];
    SiteDraw slot_draw[NP][MAXP];
    // XY: the site's angle and (cos, sin), which only this thread changes, and
    // the sweep's proposal (angle, cos, sin) formed in the draw phase
    float slot_th[NP][MAXP], slot_c[NP][MAXP], slot_s[NP][MAXP];
    float slot_pr[NP][MAXP], slot_pc[NP][MAXP], slot_ps[NP][MAXP], slot_dc[NP][MAXP], slot_ds[NP][MAXP];
    auto slot_global = [&](int c, int ri) {
        const int lr = 1 + 2 * ri + (c >> 1);
        return rep + (size_t)((row0 + lr - 1) % L) * L + 2 * tcol + (c & 1);
    };
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
        for (int m = 0; m < MAXP; m++) {
            int c = 0, ri = 0;
            const bool on = trow < rows_per_pass && pass_item(p, trow + m * rows_per_pass, HH, c, ri);
            const int pc = c & 1, base = (1 + 2 * ri + (c >> 1)) * L;
            slot_on[p][m] = on;
            slot_own[p][m] = on && !((c < 2) & (ri == HH));  // the ghost row's moves are its owner's
            slot_self[p][m] = base + pc * half + tcol;
            slot_lf[p][m] = pc ? base + tcol : base + half + (tcol == 0 ? half - 1 : tcol - 1);
            slot_rt[p][m] = pc ? base + (tcol == half - 1 ? 0 : tcol + 1) : base + half + tcol;
            slot_s0[p][m] = slot_s1[p][m] = 0;
            if (on) {
                const size_t gk = slot_global(c, ri);
                slot_s0[p][m] = a.rng[2 * gk];
                slot_s1[p][m] = a.rng[2 * gk + 1];
            }
            slot_draw[p][m] = SiteDraw{0.0f, 0.0f};
            slot_pr[p][m] = slot_pc[p][m] = slot_ps[p][m] = slot_dc[p][m] = slot_ds[p][m] = 0.0f;
            if (p == 2 || p == 3) {
                const int bnd = (on && c < 2 && ri == 0) ? 1 : (on && c < 2 && ri == HH) ? 2 : 0;
                slot_bnd[p - 2][m] = bnd;
                const int j = 2 * tcol + (c & 1);
                slot_hw[p - 2][m] = bnd == 1 ? (int)(((((size_t)r * S + s_up) * 2 + 1) * 2) * L) + j
                                             : (int)(((((size_t)r * S + s_dn) * 2 + 0) * 2) * L) + j;
                slot_pre[p - 2][m] = 0;
            }
        }
    // ---- load the strip's state: rows 0..H+2 (global rows row0-1 .. row0+H+1),
    // the loads of kLoadBatch sites per thread issued before any is used
    constexpr int kLoadBatch = 8;
    for (int base = tid; base < (H + 3) * L; base += NT * kLoadBatch) {
        float thv[kLoadBatch];
        int spv[kLoadBatch];
#pragma unroll
        for (int b = 0; b < kLoadBatch; b++) {
            const int idx = base + b * NT;
            if (idx >= (H + 3) * L) continue;
            const int lr = idx / L, j = idx - lr * L;
            const size_t gk = rep + (size_t)((row0 + lr - 1 + L) % L) * L + j;
            if (MODEL == MODEL_ISING)
                spv[b] = reinterpret_cast<const int8_t *>(a.state)[gk];
            else
                thv[b] = reinterpret_cast<const float *>(a.state)[gk];
        }
#pragma unroll
        for (int b = 0; b < kLoadBatch; b++) {
            const int idx = base + b * NT;
            if (idx >= (H + 3) * L) continue;
            const int lr = idx / L, j = idx - lr * L;
            if (MODEL == MODEL_ISING) {
                lsp[lds_index(lr, j, L)] = (int8_t)spv[b];
            } else {
                const float th = thv[b];
                float sv, cv;
                fast_sincosf(th, sv, cv);
                lcs[lds_index(lr, j, L)] = make_float2(cv, sv);
                if (lr >= 1 && lr <= H + 1) lth[lds_index(lr - 1, j, L)] = th;
            }
        }
    }
    // the PT state into LDS: ladder[w2t[w]] = betas[w] (w2t is a permutation),
    // the PT streams, the round-trip record and the widths
    if (use_pt) {
        if (tid < R) ladder[pt_walker_t] = pt_beta;
        for (int k = tid + NT; k < R; k += NT) ladder[a.w2t[k]] = a.betas[k];
        if (tid < 2 * (R - 1)) pt_g[tid] = pt_stream;
        for (int k = tid + NT; k < 2 * (R - 1); k += NT) pt_g[k] = a.pt_rng[k];
        if (s == 0 && tid < 5) my_trips[tid] = pt_trip;
    }
    if (MODEL == MODEL_XY_GAUSS) {
        if (tid < R) sig_tab[tid] = pt_sigma;
        for (int k = tid + NT; k < R; k += NT) sig_tab[k] = a.sigmas[k];
    }
    if (MODEL == MODEL_ISING && tid < 5) thr[tid] = ising_threshold(4 * tid - 8, my_beta, a.J);
    bool exchanged = false;  // my_e holds this walker's energy at this launch's last exchange
    double my_e = 0.0;       // (thread 0)

    if (MODEL == MODEL_XY_GAUSS) normal_table_k_to_lds(q_tab);
    if (tid == 0) s_err = 0;
    __syncthreads();
    float sigma = MODEL == MODEL_XY_GAUSS ? sig_tab[my_t] : 0.0f;  // this walker's proposal width
    if (MODEL != MODEL_ISING) {
#pragma unroll
        for (int p = 0; p < NP; p++)
#pragma unroll
            for (int m = 0; m < MAXP; m++) {
                const int k = slot_on[p][m] ? slot_self[p][m] : L;
                slot_th[p][m] = lth[k - L];
                const float2 q = lcs[k];
                slot_c[p][m] = q.x;
                slot_s[p][m] = q.y;
            }
    }

    int err_seen = 0;
    // The received word of boundary slot (p, m) for sweep n (tag tag0 + n):
    // the early load when it carries the tag, else polled.  A wait that
    // expires records the watchdog error and sets s_err; the sweep loop
    // leaves at the end of the sweep (see the sweep schedule below).
    auto received = [&](const int p, const int m, const int n) -> uint32_t {
        const unsigned tag = a.recv_tag0 + (unsigned)n;
        const uint64_t w = slot_pre[p - 2][m];
        uint32_t payload = (uint32_t)w;
        if ((unsigned)(w >> 32) != tag &&
            !poll_word(a.halo + slot_hw[p - 2][m] + ((n - 1) & 1) * L, tag, payload, a.error, a.watchdog)) {
            atomicOr(a.error, BATCH_ERR_WATCHDOG);
            s_err = 1;
            payload = 0;
        }
        return payload;
    };
    // The sites of pass p with this sweep's draws; `recv`: boundary sites
    // take their row-0 or row-(H+2) neighbour from the received word (S > 1,
    // after the launch's first sweep), else from LDS.
    auto apply = [&](const int p, const float beta, const bool recv, const int n) {
#pragma unroll
        for (int m = 0; m < MAXP; m++) {
            if (!slot_on[p][m]) continue;
            const bool bnd = (p == 2 || p == 3) && recv && slot_bnd[p == 3 ? 1 : 0][m] != 0;
            const bool bnd_up = bnd && slot_bnd[p == 3 ? 1 : 0][m] == 1;
            if (MODEL == MODEL_ISING) {
                const int self = slot_self[p][m], up = self - L, dn = self + L;
                const int sv = lsp[self];
                int s_up_v = lsp[up], s_dn_v = lsp[dn];
                if (bnd) {
                    const int rv = (int)(int8_t)(uint8_t)received(p, m, n);
                    if (bnd_up) s_up_v = rv; else s_dn_v = rv;
                }
                const int nb = s_up_v + s_dn_v + (int)lsp[slot_lf[p][m]] + (int)lsp[slot_rt[p][m]];
                if (slot_draw[p][m].a < thr[(2 * sv * nb + 8) >> 2]) {  // site_apply's Ising rule
                    lsp[self] = (int8_t)(-sv);
                    n_acc += slot_own[p][m];
                }
            } else {
                const int self = slot_self[p][m], lf = slot_lf[p][m], rt = slot_rt[p][m];
                const int up = self - L, dn = self + L;
                // All four neighbour reads are issued before the first add:
                // left to itself the scheduler interleaves each read with a
                // wait and an add, LDS round trips in a row per pass
                // (measured on MI355X: 124.5 -> 144 site-updates/ns).
                const float2 nu = lcs[up], nd = lcs[dn], nl = lcs[lf], nr = lcs[rt];
                float cu = nu.x, cd = nd.x;
                const float cl = nl.x, cr = nr.x;
                float su = nu.y, sd = nd.y;
                const float sl = nl.y, sr = nr.y;
                if (bnd) {
                    float rs, rc;
                    fast_sincosf(__uint_as_float(received(p, m, n)), rs, rc);
                    if (bnd_up) {
                        cu = rc;
                        su = rs;
                    } else {
                        cd = rc;
                        sd = rs;
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
                const float Cs = ((cu + cd) + cl) + cr;
                const float Ss = ((su + sd) + sl) + sr;
                float t = slot_dc[p][m] * Cs;
                t = fmaf(slot_ds[p][m], Ss, t);
                const float dE = -J * t;
                const float w = MODEL == MODEL_XY_GAUSS ? accept_weight_k21(-beta * dE) : accept_weight(-beta * dE);
                if (slot_draw[p][m].b < w) {
                    const float pr = slot_pr[p][m];
                    lth[up] = pr;
                    lcs[self] = make_float2(slot_pc[p][m], slot_ps[p][m]);
                    slot_th[p][m] = pr;
                    slot_c[p][m] = slot_pc[p][m];
                    slot_s[p][m] = slot_ps[p][m];
                    n_acc += slot_own[p][m];
                }
            }
        }
    };
    auto halo_ptr = [&](int ss, int which, int par) {
        return a.halo + ((((size_t)r * S + ss) * 2 + which) * 2 + par) * L;
    };
    // Rows 0 and H+2 for sweep n are the bottom row of strip s-1 and row 2 of
    // strip s+1 as of the end of sweep n-1 (tag tag0 + n).  The boundary
    // slots load their words early, in the draw phase (after the draws of
    // the first kPrefetchPass passes' sites), and use them in passes 2 and 3;
    // only a word that had not arrived yet is polled again.
    auto prefetch = [&](int n) {
        if (S == 1 || n == 0) return;
#pragma unroll
        for (int p = 2; p < 4; p++)
#pragma unroll
            for (int m = 0; m < MAXP; m++)
                if (slot_bnd[p - 2][m])
                    slot_pre[p - 2][m] = __hip_atomic_load(a.halo + slot_hw[p - 2][m] + ((n - 1) & 1) * L,
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    // With one strip, rows 0 and H+2 are this strip's rows H and 2, which
    // still hold sweep n-1 before pass 2 (their colour-2 and colour-3 sites
    // are updated in passes 2 to 4): copied in LDS before pass 2.
    auto copy_single = [&]() {
        for (int k = tid; k < 2 * L; k += NT) {
            const bool top = k < L;
            const int j = top ? k : k - L;
            const int dst = lds_index(top ? 0 : H + 2, j, L), src = lds_index(top ? H : 2, j, L);
            if (MODEL == MODEL_ISING) {
                lsp[dst] = lsp[src];
            } else {
                lcs[dst] = lcs[src];
            }
        }
        __syncthreads();
    };
    // After the colour-3 pass of sweep n: publish rows 2 and H (tag tag0 + n + 1).
    // Mirrors that pass's thread map, so each thread publishes the colour-3
    // site it just wrote and its colour-2 partner (written before the last
    // barrier): no extra barrier.
    auto publish = [&](int n) {
        if (S == 1 || trow >= rows_per_pass) return;
        const int par = n & 1;
        for (int q = trow; q < HH; q += rows_per_pass) {
            if (q != 0 && q != HH - 1) continue;
            const int lr = 2 + 2 * q;
            uint64_t *dst = halo_ptr(s, q == 0 ? 0 : 1, par);
            // both columns read before the first store
            uint32_t payload[2];
#pragma unroll
            for (int pp = 0; pp < 2; pp++) {
                const int j = 2 * tcol + pp;
                payload[pp] = MODEL == MODEL_ISING ? (uint32_t)(uint8_t)lsp[lds_index(lr, j, L)]
                                                   : __float_as_uint(lth[lds_index(lr - 1, j, L)]);
            }
#pragma unroll
            for (int pp = 0; pp < 2; pp++) publish_word(dst + 2 * tcol + pp, a.tag0 + (unsigned)(n + 1), payload[pp]);
        }
    };

    // Add this workgroup's accepted moves to the counter of its temperature.
    // Called by every thread (contains barriers).
    auto flush_accepted = [&]() {
        if (!a.accepted) return;
        int x = n_acc;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
        if ((tid & 63) == 0) red_i[tid >> 6] = x;
        __syncthreads();
        if (tid == 0) {
            int total = 0;
#pragma unroll
            for (int w = 0; w < NT / 64; w++) total += red_i[w];
            if (total) atomicAdd(&a.accepted[my_t], (unsigned long long)total);
        }
        __syncthreads();
        n_acc = 0;
    };

    // One sweep = colours 0, 1, 2, 3 in the reference order.  A site's update
    // reads only neighbouring colours, so any order in which
    //   c1(row) follows c0(row), c2(row) follows c0(rows +-1),
    //   c3(row) follows c2(row) and c1(rows +-1),
    //   and each site's update precedes its neighbours' next update
    // gives identical results.  Per sweep n:
    //   D   every site's draws for the sweep, from its stream (registers only)
    //   P1  c0 of interior 0/1 rows (local 3..H-1)
    //   P2  c1 of the same rows
    //   P3  c0 of rows 1 and H+1 (ghost), with rows 0 and H+2 (sweep n-1)
    //       from the received words; c2 of local rows 4..H-2
    //   P4  c1 of rows 1 and H+1, likewise; c2 of rows 2 and H
    //   P5  c3 of all 2/3 rows; publish rows 2 and H
    // P2 and P3 read nothing the other writes (P3 reads colours 1 and 3 of
    // sweep n-1 and colour 0 of P1), so no barrier separates them.  P5 of
    // sweep n and P1 of sweep n+1 touch disjoint sites and P2 starts after a
    // barrier, so P5 needs no barrier of its own: three barriers per sweep.
    long long x_next = a.x_next, x_index = a.x_index, s_next = a.s_next, s_index = a.s_index;
    // Every load of the set-up (the RNG streams above all) completes here.
    // Otherwise the compiler may wait for them inside the sweep loop, before
    // each site's first draw, and such a wait also holds the wave until the
    // loop's own memory operations in flight (the previous sweep's halo
    // stores, the early halo load) complete: measured on MI355X, 146 against
    // 162 site-updates/ns (XY L=128 R=32 Gaussian, exchanges every 10 sweeps).
    __builtin_amdgcn_s_waitcnt(0);
    for (int n = 0; n < a.nsweeps; n++) {
        const long long sweep_global = a.sweep0 + n + 1;
        const float beta = my_beta;
#pragma unroll
        for (int p = 0; p < NP; p++) {
            if (p == kPrefetchPass) prefetch(n);
#pragma unroll
            for (int m = 0; m < MAXP; m++)
                if (slot_on[p][m]) {
                    Rng g;
                    g.s0 = slot_s0[p][m];
                    g.s1 = slot_s1[p][m];
                    SiteDraw d;
                    if (MODEL == MODEL_XY_GAUSS) {
                        const GaussDraw gd = gauss_draw_k(g.next(), q_tab);  // z; v as k3 (see accept_weight_k21)
                        d.a = gd.z;
                        d.b = gd.v;
                    } else {
                        d = site_draw<MODEL>(g, q_tab);
                    }
                    slot_draw[p][m] = d;
                    slot_s0[p][m] = g.s0;
                    slot_s1[p][m] = g.s1;
                    if (MODEL != MODEL_ISING) {
                        const float th = slot_th[p][m];
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
                        slot_pr[p][m] = pr;
                        slot_pc[p][m] = cp;
                        slot_ps[p][m] = sp;
                        slot_dc[p][m] = cp - slot_c[p][m];
                        slot_ds[p][m] = sp - slot_s[p][m];
                    }
                }
        }
        const bool recv = S > 1 && n > 0;
        apply(0, beta, false, n);
        __syncthreads();
        apply(1, beta, false, n);
        if (S == 1) copy_single();
        apply(2, beta, recv, n);
        __syncthreads();
        apply(3, beta, recv, n);
        __syncthreads();
        err_seen = s_err;
        apply(4, beta, false, n);
        publish(n);
        if (err_seen) goto done;

        const bool do_exchange = n == x_next;
        const bool do_sample = n == s_next;
        if (!do_exchange && !do_sample) continue;
        __syncthreads();  // the observables read every row

        // ---- strip observables: right and down bonds (the row below row H is
        // the ghost row H+1, current for this sweep)
        long long ef = 0;  // integer energy sum of the strip
        double v[kChannels];
        double nacc = 0.0;  // the workgroup's accepted moves travel with the sums
        if (do_sample) {
            double m[4] = {0.0, 0.0, 0.0, 0.0};
            int oi = obs_i, oj = obs_j;
#pragma unroll 2
            for (int idx = tid; idx < H * L; idx += NT) {
                site_observables<MODEL>(view, 1 + oi, oj, L, true, ef, m);
                oi += obs_di;
                oj += obs_dj;
                if (oj >= L) {
                    oj -= L;
                    oi += 1;
                }
            }
            double w[5] = {m[0], m[1], m[2], m[3], (double)n_acc};
            block_sum<5, NT>(ef, w, red_e, red);
#pragma unroll
            for (int c = 0; c < 4; c++) v[c + 1] = w[c];
            nacc = w[4];
        } else {
            // The energy alone, for an exchange: each thread adds the bonds of
            // the own sites of its update slots (every own site once), with
            // the site's (cos, sin) from its registers and the neighbours'
            // LDS positions the slots already hold -- the same per-site terms
            // as site_observables, so the same integer sum.
            // Every slot's neighbour reads are issued before the first sum
            // (the positions of a slot that is off are valid LDS positions
            // too): with a branch per slot each slot took two LDS round trips
            // of its own.
            if (MODEL == MODEL_ISING) {
#pragma unroll
                for (int p = 0; p < NP; p++)
#pragma unroll
                    for (int m = 0; m < MAXP; m++) {
                        if (!slot_own[p][m]) continue;
                        const int self = slot_self[p][m], dn = self + L, rt = slot_rt[p][m];
                        const int sv = lsp[self];
                        ef += sv * ((int)lsp[dn] + (int)lsp[rt]);
                    }
            } else {
                float dc[NP][MAXP], dsn[NP][MAXP], rc[NP][MAXP], rsn[NP][MAXP];
#pragma unroll
                for (int p = 0; p < NP; p++)
#pragma unroll
                    for (int m = 0; m < MAXP; m++) {
                        const int dn = slot_self[p][m] + L, rt = slot_rt[p][m];
                        const float2 qd = lcs[dn], qr = lcs[rt];
                        dc[p][m] = qd.x;
                        dsn[p][m] = qd.y;
                        rc[p][m] = qr.x;
                        rsn[p][m] = qr.y;
                    }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int p = 0; p < NP; p++)
#pragma unroll
                    for (int m = 0; m < MAXP; m++) {
                        const float cs = slot_c[p][m], ss = slot_s[p][m];
                        const float bxc = fmaf(ss, dsn[p][m], cs * dc[p][m]);
                        const float byc = fmaf(ss, rsn[p][m], cs * rc[p][m]);
                        const long long b = bond_fixed(bxc + byc);
                        ef += slot_own[p][m] ? b : 0;
                    }
            }
            // the workgroup's totals in thread 0: wavefront sums by DPP, then
            // the NT / 64 wavefront totals (integers: exact in any order)
            const long long we = wave_sum_i64(ef), wa = wave_sum_i64((long long)n_acc);
            if ((tid & 63) == 0) {
                red_e[tid >> 6] = we;
                red_i[tid >> 6] = (int)wa;
            }
            __syncthreads();
            if (tid == 0) {
                long long te = 0;
                int ta = 0;
#pragma unroll
                for (int w = 0; w < NT / 64; w++) {
                    te += red_e[w];
                    ta += red_i[w];
                }
                ef = te;
                nacc = (double)ta;
            }
        }
        v[0] = energy_from_fixed(MODEL, J, ef);
        // moves so far belong to the current temperature
        if (tid == 0 && a.accepted && nacc > 0.0) atomicAdd(&a.accepted[my_t], (unsigned long long)nacc);
        n_acc = 0;

        if (do_exchange) {
            exchanged = true;
            const long long xi = x_index++;
            x_next += a.exchange_interval;
            const unsigned tag = a.tag0 + (unsigned)(n + 1);  // this sweep's tag
            // energy words by ladder temperature, a ring of four exchanges
            uint64_t *ring = a.energy + (size_t)(xi & 3) * R * S * 2;
            if (tid == 0) {
                uint64_t *dst = ring + ((size_t)my_t * S + s) * 2;
                publish_word(dst, tag, (uint32_t)(unsigned long long)ef);
                publish_word(dst + 1, tag, (uint32_t)((unsigned long long)ef >> 32));
            }
            // this walker's pair in the matching of parity `parity`: left
            // temperatures parity, parity + 2, ... below R - 1
            // (core/parallel_tempering.py:44-69)
            const int rel = my_t - parity;
            const int partner = (rel >= 0 && !(rel & 1) && my_t < R - 1) ? my_t + 1
                                : (rel > 0 && (rel & 1)) ? my_t - 1 : -1;
            const int my_left = partner < 0 ? -1 : (partner < my_t ? partner : my_t);
            // Every attempted pair's stream takes one draw (energies are
            // integer sums here, always finite).  The threads of the second
            // wavefront take the draws of the other pairs while the third
            // wavefront (the deciding one) takes this pair's draw, gathers
            // the strips of this walker and of its partner, adds them with
            // DPP sums and decides.  The draws do not depend on the energies,
            // so they run while the energy words are in flight.
            for (int left = parity + 2 * ((tid + NT - 64) % NT); left < R - 1; left += 2 * NT) {
                if (left == my_left) continue;
                Rng g;
                g.s0 = pt_g[2 * left];
                g.s1 = pt_g[2 * left + 1];
                (void)g.next();
                pt_g[2 * left] = g.s0;
                pt_g[2 * left + 1] = g.s1;
            }
            if ((tid >> 6) == 2) {
                const int lane = tid & 63;
                double u = 0.0, logu = 0.0;
                if (my_left >= 0) {
                    Rng g;
                    g.s0 = pt_g[2 * my_left];
                    g.s1 = pt_g[2 * my_left + 1];
                    u = g.uniform_f64();
                    logu = u == 0.0 ? 0.0 : ::log(u);
                    if (lane == 0) {
                        pt_g[2 * my_left] = g.s0;
                        pt_g[2 * my_left + 1] = g.s1;
                    }
                }
                // word k: strip k / 2 (low half for even k) of this walker
                // (k < 2 S), then of the partner
                const int nw = partner >= 0 ? 4 * S : 2 * S;
                int bad = 0;
                long long e_own = 0, e_par = 0;
                for (int k0 = 0; k0 < nw; k0 += 64) {
                    const int k = k0 + lane;
                    uint32_t payload = 0;
                    if (k < nw) {
                        const int t = k < 2 * S ? my_t : partner;
                        if (!poll_word(ring + (size_t)t * S * 2 + (k % (2 * S)), tag, payload, a.error, a.watchdog))
                            bad = 1;
                    }
                    // the high half from the next lane (quad permutation 1, 1, 3, 3)
                    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)payload, 0xF5, 0xf, 0xf, false);
                    const long long w = (k & 1) ? 0 : (long long)(((unsigned long long)hi << 32) | payload);
                    e_own += wave_sum_i64(k < 2 * S ? w : 0);
                    e_par += wave_sum_i64(k < 2 * S ? 0 : w);
                }
                if (lane == 0) {
                    // this walker's pair decides (core/parallel_tempering.py:44-69)
                    const double e0 = energy_from_fixed(MODEL, J, e_own);
                    int acc = 0;
                    if (my_left >= 0) {
                        const double e1 = energy_from_fixed(MODEL, J, e_par);
                        const double ea = my_t == my_left ? e0 : e1, eb = my_t == my_left ? e1 : e0;
                        const double lr = ((double)ladder[my_left] - (double)ladder[my_left + 1]) * (ea - eb);
                        acc = (lr >= 0.0 || u == 0.0 || logu < lr) ? 1 : 0;
                    }
                    x_acc = acc;
                    x_e = e0;
                }
                if (__any(bad)) {
                    if (lane == 0) atomicOr(a.error, BATCH_ERR_WATCHDOG);
                    s_err = 1;
                }
            }
            __syncthreads();
            if (s_err) goto done;
            const bool acc = x_acc != 0;
            if (tid == 0) my_e = x_e;
            if (s == 0 && tid == 0 && my_left == my_t) {  // the left walker's strip 0 counts the pair
                atomicAdd(reinterpret_cast<unsigned long long *>(a.counts) + my_left, 1ull);
                if (acc) atomicAdd(reinterpret_cast<unsigned long long *>(a.counts) + (R - 1) + my_left, 1ull);
            }
            if (acc) my_t = partner;
            my_beta = ladder[my_t];
            if (MODEL == MODEL_XY_GAUSS) sigma = sig_tab[my_t];
            // this walker's round trips (core/parallel_tempering.py:71-91)
            if (s == 0 && tid == 0 && (my_t == 0 || my_t == R - 1)) {
                const long long origin = my_trips[0];
                if (origin < 0) {
                    my_trips[0] = my_t;
                    my_trips[2] = sweep_global;
                } else if (my_t != origin) {
                    my_trips[1] = 1;
                } else if (my_trips[1]) {
                    my_trips[3] += 1;
                    my_trips[4] += sweep_global - my_trips[2];
                    my_trips[2] = sweep_global;
                    my_trips[1] = 0;
                }
            }
            if (R > 2) parity = 1 - parity;
            if (MODEL == MODEL_ISING) {
                if (tid < 5) thr[tid] = ising_threshold(4 * tid - 8, my_beta, a.J);
                __syncthreads();
            }
        }

        if (do_sample) {
            const long long idx = s_index++;
            s_next += a.sample_interval;
            if (idx < a.max_samples) {
                if (tid == 0) {
                    double *dst = a.sample_partials + (((size_t)idx * R + r) * S + s) * kChannels;
#pragma unroll
                    for (int c = 0; c < kChannels; c++) dst[c] = v[c];
                }
                if (s == 0 && tid == 0) a.sample_w2t[(size_t)idx * R + r] = my_t;
            }
        }
    }

done:
    flush_accepted();
    // the PT state after the last exchange: betas, maps, PT streams and parity
    // (read by every workgroup at its start) to the staging area, the rest --
    // read only by the workgroup that writes it -- in place
    const BatchStage stage = batch_stage(a.energy, R, S);
    if (use_pt) {
        if (s == 0 && tid == 0) {  // this walker's temperature, beta, energy and record
            stage.betas[r] = my_beta;
            stage.w2t[r] = my_t;
            stage.t2w[my_t] = r;
            if (exchanged) a.energies[r] = my_e;
            for (int row = 0; row < 5; row++) a.trips[row * R + r] = my_trips[row];
        }
        if (blockIdx.x == 0)
            for (int k = tid; k < 2 * (R - 1); k += NT) stage.pt_rng[k] = pt_g[k];
    }
    // ---- write back own rows' state and the streams of owned sites
    __syncthreads();
    for (int idx = tid; idx < H * L; idx += NT) {
        const int lr = 1 + idx / L, j = idx - (lr - 1) * L;
        const size_t gk = rep + (size_t)(row0 + lr - 1) * L + j;
        if (MODEL == MODEL_ISING)
            reinterpret_cast<int8_t *>(a.state)[gk] = lsp[lds_index(lr, j, L)];
        else
            reinterpret_cast<float *>(a.state)[gk] = lth[lds_index(lr - 1, j, L)];
    }
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
        for (int m = 0; m < MAXP; m++)
            if (slot_own[p][m]) {
                int c = 0, ri = 0;
                pass_item(p, trow + m * rows_per_pass, HH, c, ri);
                const size_t gk = slot_global(c, ri);
                a.rng[2 * gk] = slot_s0[p][m];
                a.rng[2 * gk + 1] = slot_s1[p][m];
            }
    if (blockIdx.x == 0 && tid == 0 && use_pt) *stage.parity = parity;
}

// Queued behind a launch on its stream: the staged PT state in place (after a
// launch with exchanges), and the watchdog code into the caller's pinned host
// word (error_host, device-accessible; NULL: none) -- a store of this kernel
// instead of a copy operation of its own.
__global__ void batch_commit_kernel(const uint64_t *energy, int R, int S, float *betas, int32_t *w2t,
                                    int32_t *t2w, uint64_t *pt_rng, int *parity, int exchanges, const int *error,
                                    int *error_host) {
    if (exchanges) {
        const BatchStage stage = batch_stage(const_cast<uint64_t *>(energy), R, S);
        for (int k = threadIdx.x; k < R; k += blockDim.x) {
            betas[k] = stage.betas[k];
            w2t[k] = stage.w2t[k];
            t2w[k] = stage.t2w[k];
        }
        for (int k = threadIdx.x; k < 2 * (R - 1); k += blockDim.x) pt_rng[k] = stage.pt_rng[k];
        if (threadIdx.x == 0) *parity = *stage.parity;
    }
    if (error_host && threadIdx.x == 0) *error_host = *error;
}

// Zeroes the workspace's halo and energy words and the error word.  A kernel
// rather than memset nodes, so a launch can be captured into a graph.
__global__ void batch_reset_kernel(uint64_t *halo, size_t nhalo, uint64_t *energy, size_t nenergy, int *error) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < nhalo; k += stride) halo[k] = 0;
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < nenergy; k += stride) energy[k] = 0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *error = 0;
}

// Test hooks (mc_batch_test_watchdog): receives that wait for tags no
// workgroup publishes, and a shorter watchdog; both 0 in normal use.
static int g_batch_test_unreachable = 0;
// Launch mode of mc_batch_run (mc_batch_set_cooperative): a plain dispatch
// after an occupancy check on the host (the default: on MI355X it starts
// 20-30 us sooner than a cooperative launch of the same grid, measured on
// 200-sweep launches of XY L=128 R=32), or a cooperative launch, which the
// runtime admits only when every workgroup is co-resident.  Either way every
// wait is bounded, so a grid that is not co-resident ends in the watchdog
// error rather than a hang.
static bool g_batch_cooperative = false;

// co-resident workgroups per (device, kernel, workgroup size, dynamic LDS)
struct Capacity {
    int dev;
    const void *kern;
    int threads;
    size_t lds;
    long long capacity;
};
static std::mutex g_capacity_lock;
static std::vector<Capacity> g_capacity;
static unsigned g_batch_test_watchdog = 0;

static size_t halo_words(int L, int R, int H) { return (size_t)R * (L / H) * 2 * 2 * L; }
// Bound of the workspace clock (tags are 32-bit); receives of the test hook
// wait for tags this far above any published one
static constexpr long long kMaxTag = 0x3fffffffLL;
static constexpr unsigned kUnreachableTag = 0x40000000u;
static size_t energy_words(int L, int R, int H) { return (size_t)4 * R * (L / H) * 2 + batch_stage_words(R); }

static size_t batch_lds(int model, int L, int R, int H, int exchange) {
    (void)exchange;
    const size_t pt_bytes = (size_t)16 * (R - 1) + (size_t)8 * R;
    return ((strip_bytes(model, L, H, false) + 15) & ~(size_t)15) + pt_bytes + 64;
}

// Items per pass and thread: the largest pass has max(H/2, 4) colour rows,
// spread over threads / (L/2) colour-row slots
static int batch_items(int L, int H, int threads) {
    const int rpp = threads / (L / 2);
    const int most = std::max(H / 2, 4);
    return (most + rpp - 1) / rpp;
}

template <int NT, int MAXP>
static const void *kernel_ptr_nm(int model) {
    if (model == MODEL_ISING) return reinterpret_cast<const void *>(&batch_kernel<MODEL_ISING, NT, MAXP>);
    if (model == MODEL_XY_FAST) return reinterpret_cast<const void *>(&batch_kernel<MODEL_XY_FAST, NT, MAXP>);
    return reinterpret_cast<const void *>(&batch_kernel<MODEL_XY_GAUSS, NT, MAXP>);
}

static const void *kernel_ptr(int model, int threads, int items) {
    if (threads == 512) return items <= 1 ? kernel_ptr_nm<512, 1>(model) : kernel_ptr_nm<512, 2>(model);
    return items <= 1 ? kernel_ptr_nm<256, 1>(model) : kernel_ptr_nm<256, 2>(model);
}

static int check_strips(int model, int L, int R, int H) {
    MC_REQUIRE(model == MODEL_ISING || model == MODEL_XY_FAST || model == MODEL_XY_GAUSS,
               "unknown batch model %d", model);
    MC_REQUIRE(L >= 4 && L % 2 == 0, "L must be even and >= 4, got %d", L);
    MC_REQUIRE(H >= 4 && H % 2 == 0 && L % H == 0, "strip height %d must be even, >= 4 and divide L=%d",
               H, L);
    MC_REQUIRE(R >= 1, "need at least one replica");
    return MC_OK;
}

static int check_shape(int model, int L, int R, int H, int threads) {
    MC_REQUIRE(threads == 256 || threads == 512, "workgroup size must be 256 or 512, got %d", threads);
    int rc = check_strips(model, L, R, H);
    if (rc) return rc;
    MC_REQUIRE(L / 2 <= threads, "L=%d is too wide for a %d-thread workgroup (L/2 <= threads)", L, threads);
    MC_REQUIRE(batch_items(L, H, threads) <= 2,
               "strips of %d rows of L=%d need more than two sites per pass and thread of %d", H, L, threads);
    MC_REQUIRE(R >= 1, "need at least one replica");
    return MC_OK;
}

// The arguments of a run that stay the same from one launch to the next,
// checked once, with the kernel, its LDS and the co-residency check resolved.
struct BatchPlan {
    void *stream = nullptr;
    BatchArgs a{};  // every field but the per-launch ones
    const void *kern = nullptr;
    size_t lds = 0;
    int threads = 0;
    long long *clock = nullptr;
    int *error_host = nullptr;
};

static int batch_prepare(BatchPlan &p, void *stream, int model, int L, int R, int H, int threads, void *state,
                         uint64_t *rng, float *betas, float J, float delta, const float *sigmas,
                         unsigned long long *accepted, int exchange_interval, int *parity, double *energies,
                         int32_t *w2t, int32_t *t2w, uint64_t *pt_rng, long long *counts, long long *trips,
                         uint64_t *halo, uint64_t *energy, int *error, long long *clock, int *error_host) {
    int rc = check_shape(model, L, R, H, threads);
    if (rc) return rc;
    MC_REQUIRE(exchange_interval >= 0, "exchange_interval must be nonnegative");
    MC_REQUIRE(exchange_interval == 0 || R >= 2, "exchanges need at least two replicas");
    MC_REQUIRE(model != MODEL_XY_GAUSS || sigmas, "the Gaussian XY model needs per-temperature sigmas");
    const int S = L / H;
    const size_t lds = batch_lds(model, L, R, H, exchange_interval > 0);
    MC_REQUIRE(lds <= 160 * 1024, "batch workgroup needs %zu bytes of LDS", lds);

    BatchArgs &a = p.a;
    a.model = model; a.L = L; a.R = R; a.H = H; a.S = S;
    a.state = state; a.rng = rng; a.betas = betas; a.J = J; a.delta = delta;
    a.sigmas = sigmas; a.accepted = accepted;
    a.exchange_interval = exchange_interval; a.parity = parity; a.energies = energies;
    a.w2t = w2t; a.t2w = t2w; a.pt_rng = pt_rng; a.counts = counts; a.trips = trips;
    a.halo = halo; a.energy = energy; a.error = error;

    const void *kern = kernel_ptr(model, threads, batch_items(L, H, threads));
    if (lds > 64 * 1024) MC_TRY_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    // a plain dispatch places every workgroup at once only for a grid the
    // occupancy check admits (a cooperative launch checks the same)
    int dev = 0;
    MC_TRY_HIP(hipGetDevice(&dev));
    long long capacity = -1;
    {
        std::lock_guard<std::mutex> guard(g_capacity_lock);
        for (const auto &e : g_capacity)
            if (e.dev == dev && e.kern == kern && e.threads == threads && e.lds == lds) capacity = e.capacity;
    }
    if (capacity < 0) {
        int per_cu = 0, cus = 0;
        MC_TRY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds));
        MC_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        capacity = (long long)per_cu * cus;
        std::lock_guard<std::mutex> guard(g_capacity_lock);
        g_capacity.push_back(Capacity{dev, kern, threads, lds, capacity});
    }
    MC_REQUIRE((long long)R * S <= capacity, "%d workgroups cannot be co-resident (%lld on the device at %zu "
               "bytes of LDS)", R * S, capacity, lds);
    p.stream = stream;
    p.kern = kern;
    p.lds = lds;
    p.threads = threads;
    p.clock = clock;
    p.error_host = error_host;
    return MC_OK;
}

static int batch_launch(const BatchPlan &p, int nsweeps, long long sweep0, int sample_interval,
                        long long sample_first, int max_samples, double *sample_partials, int32_t *sample_w2t) {
    MC_REQUIRE(nsweeps >= 0, "nsweeps must be nonnegative");
    MC_REQUIRE(sample_interval >= 0 && max_samples >= 0, "invalid sampling arguments");
    if (nsweeps == 0) return MC_OK;
    BatchArgs a = p.a;
    const int L = a.L, R = a.R, H = a.H, S = a.S, exchange_interval = a.exchange_interval;
    a.nsweeps = nsweeps; a.sweep0 = sweep0;
    a.sample_interval = sample_interval; a.sample_first = sample_first; a.max_samples = max_samples;
    a.sample_partials = sample_partials; a.sample_w2t = sample_w2t;
    // first exchange sweep: the first global sweep g >= sweep0 + 1 with
    // g % exchange_interval == 0; first sample sweep: the first
    // g >= max(sweep0 + 1, sample_first) with (g - sample_first) % sample_interval == 0
    const long long g1 = sweep0 + 1;
    a.x_next = a.s_next = nsweeps;
    a.x_index = a.s_index = 0;
    if (exchange_interval > 0) {
        const long long first = g1 + (exchange_interval - g1 % exchange_interval) % exchange_interval;
        a.x_next = std::min<long long>(first - g1, nsweeps);
        a.x_index = first / exchange_interval;
    }
    if (sample_interval > 0) {
        const long long first =
            g1 <= sample_first ? sample_first
                               : g1 + (sample_interval - (g1 - sample_first) % sample_interval) % sample_interval;
        a.s_next = std::min<long long>(first - g1, nsweeps);
        a.s_index = (first - sample_first) / sample_interval;
    }
    a.watchdog = g_batch_test_watchdog ? g_batch_test_watchdog : kWatchdog;

    // The workspace clock: tags issued since the workspace was last zeroed.
    // Zeroed first when there is no clock, when it reads zero (a new
    // workspace, or one after an error) and before the count could overflow.
    long long *clock = p.clock;
    MC_REQUIRE((long long)nsweeps + 1 <= kMaxTag, "at most %lld sweeps per launch", kMaxTag - 1);
    const bool reset = clock == nullptr || clock[0] <= 0 || clock[0] + (long long)nsweeps + 1 > kMaxTag;
    const long long tag0 = reset ? 0 : clock[0];
    a.tag0 = (unsigned)tag0;
    a.recv_tag0 = a.tag0 + (g_batch_test_unreachable && S > 1 ? kUnreachableTag : 0u);

    hipStream_t st = as_stream(p.stream);
    if (reset) {
        const size_t nh = halo_words(L, R, H);
        const unsigned reset_blocks = (unsigned)std::min<size_t>((nh + 255) / 256, 1024);
        hipLaunchKernelGGL(batch_reset_kernel, dim3(reset_blocks), dim3(256), 0, st, a.halo, nh, a.energy,
                           energy_words(L, R, H), a.error);
        MC_TRY_HIP(hipGetLastError());
    }
    void *kargs[] = {&a};
    if (g_batch_cooperative)
        MC_TRY_HIP(hipLaunchCooperativeKernel(p.kern, dim3((unsigned)(R * S)), dim3(p.threads), kargs, p.lds, st));
    else
        MC_TRY_HIP(hipLaunchKernel(p.kern, dim3((unsigned)(R * S)), dim3(p.threads), kargs, p.lds, st));
    if (exchange_interval > 0 || p.error_host) {
        hipLaunchKernelGGL(batch_commit_kernel, dim3(1), dim3(256), 0, st, a.energy, R, S, a.betas, a.w2t, a.t2w,
                           a.pt_rng, a.parity, exchange_interval > 0 ? 1 : 0, a.error, p.error_host);
        MC_TRY_HIP(hipGetLastError());
    }
    if (clock) clock[0] = tag0 + nsweeps + 1;  // every tag of this launch is below it
    return MC_OK;
}

}  // namespace mc

using namespace mc;

extern "C" {

int mc_batch_test_watchdog(int unreachable, unsigned polls) {
    g_batch_test_unreachable = unreachable != 0;
    g_batch_test_watchdog = polls;
    return MC_OK;
}

int mc_batch_workspace_size(int model, int L, int R, int H, size_t *halo_words_out, size_t *energy_words_out) {
    int rc = check_strips(model, L, R, H);
    if (rc) return rc;
    *halo_words_out = halo_words(L, R, H);
    *energy_words_out = energy_words(L, R, H);
    return MC_OK;
}

int mc_batch_max_workgroups(int model, int L, int R, int H, int threads, int exchange, int *max_workgroups) {
    int rc = check_shape(model, L, R, H, threads);
    if (rc) return rc;
    const size_t lds = batch_lds(model, L, R, H, exchange);
    if (lds > 160 * 1024) {
        *max_workgroups = 0;
        return MC_OK;
    }
    if (lds > 64 * 1024)
        MC_TRY_HIP(hipFuncSetAttribute(kernel_ptr(model, threads, batch_items(L, H, threads)),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0, dev = 0, cus = 0;
    MC_TRY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel_ptr(model, threads, batch_items(L, H, threads)),
                                                            threads, lds));
    MC_TRY_HIP(hipGetDevice(&dev));
    MC_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    *max_workgroups = per_cu * cus;
    return MC_OK;
}

int mc_batch_run(void *stream, int model, int L, int R, int H, int threads, void *state, uint64_t *rng,
                 float *betas, float J, float delta, const float *sigmas, unsigned long long *accepted,
                 int nsweeps, long long sweep0, int exchange_interval, int *parity,
                 double *energies, int32_t *w2t, int32_t *t2w, uint64_t *pt_rng, long long *counts,
                 long long *trips, int sample_interval, long long sample_first, int max_samples,
                 double *sample_partials, int32_t *sample_w2t, uint64_t *halo, uint64_t *energy,
                 int *error, long long *clock, int *error_host) {
    BatchPlan plan;
    const int rc = batch_prepare(plan, stream, model, L, R, H, threads, state, rng, betas, J, delta, sigmas,
                                 accepted, exchange_interval, parity, energies, w2t, t2w, pt_rng, counts, trips,
                                 halo, energy, error, clock, error_host);
    if (rc) return rc;
    return batch_launch(plan, nsweeps, sweep0, sample_interval, sample_first, max_samples, sample_partials,
                        sample_w2t);
}

int mc_batch_plan_create(void *stream, int model, int L, int R, int H, int threads, void *state, uint64_t *rng,
                         float *betas, float J, float delta, const float *sigmas, unsigned long long *accepted,
                         int exchange_interval, int *parity, double *energies, int32_t *w2t, int32_t *t2w,
                         uint64_t *pt_rng, long long *counts, long long *trips, uint64_t *halo, uint64_t *energy,
                         int *error, long long *clock, int *error_host, void **plan) {
    MC_REQUIRE(plan, "plan must not be NULL");
    *plan = nullptr;
    BatchPlan *p = new (std::nothrow) BatchPlan;
    MC_REQUIRE(p, "out of host memory");
    const int rc = batch_prepare(*p, stream, model, L, R, H, threads, state, rng, betas, J, delta, sigmas, accepted,
                                 exchange_interval, parity, energies, w2t, t2w, pt_rng, counts, trips, halo, energy,
                                 error, clock, error_host);
    if (rc) {
        delete p;
        return rc;
    }
    *plan = p;
    return MC_OK;
}

int mc_batch_plan_launch(void *plan, int nsweeps, long long sweep0, int sample_interval, long long sample_first,
                         int max_samples, double *sample_partials, int32_t *sample_w2t) {
    MC_REQUIRE(plan, "plan must not be NULL");
    return batch_launch(*static_cast<const BatchPlan *>(plan), nsweeps, sweep0, sample_interval, sample_first,
                        max_samples, sample_partials, sample_w2t);
}

int mc_batch_plan_destroy(void *plan) {
    delete static_cast<BatchPlan *>(plan);
    return MC_OK;
}

int mc_batch_set_cooperative(int cooperative) {
    g_batch_cooperative = cooperative != 0;
    return MC_OK;
}

}  // extern "C"
