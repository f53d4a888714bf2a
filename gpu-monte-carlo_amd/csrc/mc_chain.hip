// Native chain of the drop-in path: one lattice size's walkers of one model,
// advanced by whole sweeps -- and, with exchanges attached, by the
// sweep / energy / exchange / walker-history schedule -- in one C call.
//
// The drop-in classes keep the reference's per-launch structure (four colour
// kernels per sweep, core/runtime.py:87-99; a fresh energy and one matching
// per exchange, core/parallel_tempering.py:245-269), but the loop over
// sweeps, colours and exchanges runs here, on the host side of the library:
// Python makes one call for any number of sweeps instead of four or more
// launches per sweep, and every launch goes to the chain's stream in order.
// A chain binds device buffers owned by its caller; it allocates nothing.
#include <string.h>

#include <mutex>

#include "mc_common.h"
#include "mc_reduce.h"

namespace mc {

// Reduction partials of mc_chain_energy_noalloc: a file-scope device array,
// one copy per GPU, held until the process ends.  Host code never uses the
// array's name as an address: each device's copy is found once with
// hipGetSymbolAddress (with that device current) and kept in
// g_noalloc_partials.  Every call on one GPU shares its copy, so calls that
// overlap on one GPU (different streams) race; the reduction writes every
// partial it reads, so nothing needs clearing between calls.
constexpr size_t kNoallocPartials = 65536;  // 512 KiB of float64: R x (blocks per replica, <= 64)
static __device__ __align__(256) double scratch[kNoallocPartials];
static std::mutex g_noalloc_lock;
static double *g_noalloc_partials[64];

static int noalloc_partials(double **out) {
    int dev = 0;
    MC_TRY_HIP(hipGetDevice(&dev));
    MC_REQUIRE(dev >= 0 && dev < 64, "device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> guard(g_noalloc_lock);
    if (!g_noalloc_partials[dev]) {
        void *p = nullptr;
        MC_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_noalloc_partials[dev] = static_cast<double *>(p);
    }
    *out = g_noalloc_partials[dev];
    return MC_OK;
}

// Routes this thread's reductions to a caller-provided buffer for one scope.
struct PartialsScope {
    PartialsScope(double *p, size_t n) { g_partials_override = PartialsOverride{p, n}; }
    ~PartialsScope() { g_partials_override = PartialsOverride{nullptr, 0}; }
};

enum ChainKind {
    CHAIN_ISING = 0,      // params J
    CHAIN_XY = 1,         // params J, delta
    CHAIN_XY_GAUSS = 2,   // params J, T; buffers table_betas, sigmas, accepted
    CHAIN_KAWAMURA = 3,   // params J, delta
    CHAIN_HE3 = 4,        // params c0..c4, K, delta, g_d
    CHAIN_HE3_GAUSS = 5,  // params c0..c4, K, rho, mass, g_d; buffers counts
};

struct Chain {
    int kind, code, L, R;
    void *state;
    uint64_t *rng;
    void *betas;
    double J, delta, K, g_d, rho, mass;
    double coef[5];
    int T;
    const void *table_betas, *sigmas;
    unsigned long long *accepted;
    int64_t *counts;
    // exchanges (mc_chain_attach_exchange)
    double *energies;
    int32_t *w2t, *t2w;
    uint64_t *pt_rng;
    int64_t *pt_counts, *trips;
};

static int chain_channels(int kind) {
    switch (kind) {
    case CHAIN_ISING: return 2;
    case CHAIN_XY:
    case CHAIN_XY_GAUSS: return 5;
    case CHAIN_KAWAMURA: return 7;
    default: return 10;
    }
}

static int chain_color(const Chain &c, void *stream, int color) {
    switch (c.kind) {
    case CHAIN_ISING:
        return mc_ising_update_color(stream, (int8_t *)c.state, c.rng, (const float *)c.betas, c.R, c.L, color,
                                     (float)c.J);
    case CHAIN_XY:
        return mc_xy_update_color(stream, c.code, c.state, c.rng, c.betas, c.R, c.L, color, c.J, c.delta);
    case CHAIN_XY_GAUSS:
        return mc_xy_gauss_update_color(stream, c.code, c.state, c.rng, c.betas, c.R, c.L, color, c.J,
                                        c.table_betas, c.sigmas, c.T, c.accepted);
    case CHAIN_KAWAMURA:
        return mc_kawamura_update_color(stream, c.code, c.state, c.rng, c.betas, c.R, c.L, color, c.J, c.delta);
    case CHAIN_HE3:
        return mc_he3_update_color(stream, c.code, c.state, c.rng, c.betas, c.R, c.L, color, c.coef, c.K,
                                   c.delta, c.g_d);
    default:
        return mc_he3_gaussian_update_color(stream, c.code, c.state, c.rng, c.betas, c.R, c.L, color, c.coef,
                                            c.K, c.rho, c.mass, c.counts, c.g_d);
    }
}

// One sweep: the row-staged sweep kernels where they measured faster, else
// the model's four colour launches in order (the same results either way).
// Kawamura float32 at L=192, R=64: 28.6 site-updates/ns with the row kernels
// against 27.0 with colour launches; float64: 8.3 against 13.4 (164 VGPRs
// and 54 KiB of LDS leave 2 workgroups per CU), so float64 keeps the colour
// launches.
static int chain_sweep(const Chain &c, void *stream) {
    if (c.kind == CHAIN_KAWAMURA && c.code == MC_F32)
        return mc_kawamura_sweep(stream, c.code, c.state, c.rng, c.betas, c.R, c.L, c.J, c.delta);
    if (c.kind == CHAIN_HE3)
        return mc_he3_sweep(stream, c.code, c.state, c.rng, c.betas, c.R, c.L, c.coef, c.K, c.delta, c.g_d);
    for (int color = 0; color < 4; color++) {
        const int rc = chain_color(c, stream, color);
        if (rc) return rc;
    }
    return MC_OK;
}

static int chain_energy(const Chain &c, void *stream, double *out) {
    switch (c.kind) {
    case CHAIN_ISING: return mc_ising_energy(stream, (const int8_t *)c.state, c.R, c.L, (float)c.J, out);
    case CHAIN_XY:
    case CHAIN_XY_GAUSS: return mc_xy_energy(stream, c.code, c.state, c.R, c.L, c.J, out);
    case CHAIN_KAWAMURA: return mc_kawamura_energy(stream, c.code, c.state, c.R, c.L, c.J, out);
    default: return mc_he3_energy(stream, c.code, c.state, c.R, c.L, c.coef, c.K, c.g_d, out);
    }
}

static int chain_measure(const Chain &c, void *stream, const int32_t *tind, const double *cached, double *out) {
    switch (c.kind) {
    case CHAIN_ISING:
        return mc_ising_measure(stream, (const int8_t *)c.state, c.R, c.L, (float)c.J, tind, cached, out);
    case CHAIN_XY:
    case CHAIN_XY_GAUSS: return mc_xy_measure(stream, c.code, c.state, c.R, c.L, c.J, tind, cached, out);
    case CHAIN_KAWAMURA:
        return mc_kawamura_measure(stream, c.code, c.state, c.R, c.L, c.J, tind, cached, out);
    default:
        return mc_he3_measure(stream, c.code, c.state, c.R, c.L, c.coef, c.K, c.g_d, tind, cached, out);
    }
}

// precision of the walker betas (Ising updates with float32 betas)
static int chain_beta_code(const Chain &c) { return c.kind == CHAIN_ISING ? MC_F32 : c.code; }

}  // namespace mc

using namespace mc;

extern "C" {

int mc_chain_create(int kind, int dtype, int L, int R, void *state, uint64_t *rng, void *betas,
                    const double *params, int nparams, void *const *buffers, int nbuffers, void **chain) {
    MC_REQUIRE(kind >= CHAIN_ISING && kind <= CHAIN_HE3_GAUSS, "unknown chain kind %d", kind);
    MC_REQUIRE(dtype == MC_F32 || dtype == MC_F64, "dtype must be MC_F32 or MC_F64");
    MC_REQUIRE(kind != CHAIN_ISING || dtype == MC_F32, "Ising chains use float32 betas");
    MC_REQUIRE(L >= 4 && L % 2 == 0, "L must be even and >= 4, got %d", L);
    MC_REQUIRE(R >= 1, "need at least one replica");
    MC_REQUIRE(state && rng && betas && chain, "state, rng, betas and chain are required");
    static const int want_params[] = {1, 2, 2, 2, 8, 9};
    static const int want_buffers[] = {0, 0, 3, 0, 0, 1};
    MC_REQUIRE(nparams == want_params[kind] && (nparams == 0 || params), "chain kind %d takes %d parameters",
               kind, want_params[kind]);
    MC_REQUIRE(nbuffers == want_buffers[kind] && (nbuffers == 0 || buffers), "chain kind %d takes %d buffers",
               kind, want_buffers[kind]);
    Chain *c = new Chain();
    memset(c, 0, sizeof(Chain));
    c->kind = kind;
    c->code = dtype;
    c->L = L;
    c->R = R;
    c->state = state;
    c->rng = rng;
    c->betas = betas;
    c->J = params[0];
    switch (kind) {
    case CHAIN_XY:
    case CHAIN_KAWAMURA:
        c->delta = params[1];
        break;
    case CHAIN_XY_GAUSS:
        c->T = (int)params[1];
        c->table_betas = buffers[0];
        c->sigmas = buffers[1];
        c->accepted = static_cast<unsigned long long *>(buffers[2]);
        break;
    case CHAIN_HE3:
    case CHAIN_HE3_GAUSS:
        for (int k = 0; k < 5; k++) c->coef[k] = params[k];
        c->J = 0.0;
        c->K = params[5];
        if (kind == CHAIN_HE3) {
            c->delta = params[6];
            c->g_d = params[7];
        } else {
            c->rho = params[6];
            c->mass = params[7];
            c->g_d = params[8];
            c->counts = static_cast<int64_t *>(buffers[0]);
        }
        break;
    default:
        break;
    }
    if (kind == CHAIN_XY_GAUSS && (c->T < 1 || !c->table_betas || !c->sigmas)) {
        delete c;
        MC_REQUIRE(false, "Gaussian XY chains need T >= 1, table betas and sigmas");
    }
    if (kind == CHAIN_HE3_GAUSS && !c->counts) {
        delete c;
        MC_REQUIRE(false, "Gaussian He3 chains need a counts buffer");
    }
    *chain = c;
    return MC_OK;
}

int mc_chain_destroy(void *chain) {
    delete static_cast<Chain *>(chain);
    return MC_OK;
}

int mc_chain_channels(void *chain, int *channels) {
    MC_REQUIRE(chain && channels, "null chain");
    *channels = chain_channels(static_cast<Chain *>(chain)->kind);
    return MC_OK;
}

int mc_chain_attach_exchange(void *chain, double *energies, int32_t *walker_to_temperature,
                             int32_t *temperature_to_walker, uint64_t *pt_rng, int64_t *counts, int64_t *trips) {
    Chain *c = static_cast<Chain *>(chain);
    MC_REQUIRE(c, "null chain");
    MC_REQUIRE(c->R >= 2, "exchanges need at least two replicas");
    MC_REQUIRE(energies && walker_to_temperature && temperature_to_walker && pt_rng && counts && trips,
               "exchanges need energies, both maps, the PT streams, counts and trips");
    c->energies = energies;
    c->w2t = walker_to_temperature;
    c->t2w = temperature_to_walker;
    c->pt_rng = pt_rng;
    c->pt_counts = counts;
    c->trips = trips;
    return MC_OK;
}

int mc_chain_sweeps(void *chain, void *stream, long long nsweeps) {
    Chain *c = static_cast<Chain *>(chain);
    MC_REQUIRE(c, "null chain");
    MC_REQUIRE(nsweeps >= 0, "nsweeps must be nonnegative");
    for (long long n = 0; n < nsweeps; n++) {
        const int rc = chain_sweep(*c, stream);
        if (rc) return rc;
    }
    return MC_OK;
}

int mc_chain_energy(void *chain, void *stream, double *out) {
    Chain *c = static_cast<Chain *>(chain);
    MC_REQUIRE(c && out, "null chain or output");
    return chain_energy(*c, stream, out);
}

int mc_chain_energy_noalloc(void *chain, void *stream, double *out) {
    Chain *c = static_cast<Chain *>(chain);
    MC_REQUIRE(c && out, "null chain or output");
    const size_t need = (size_t)c->R * reduce_blocks(c->L);
    MC_REQUIRE(need <= kNoallocPartials,
               "R=%d replicas of L=%d need %zu reduction partials; mc_chain_energy_noalloc holds at most %zu "
               "(use mc_chain_energy)", c->R, c->L, need, kNoallocPartials);
    double *partials = nullptr;
    const int rc = noalloc_partials(&partials);
    if (rc) return rc;
    PartialsScope scope(partials, kNoallocPartials);
    return chain_energy(*c, stream, out);
}

int mc_chain_measure(void *chain, void *stream, const int32_t *temperature_indices, const double *cached_energies,
                     double *out) {
    Chain *c = static_cast<Chain *>(chain);
    MC_REQUIRE(c && out, "null chain or output");
    return chain_measure(*c, stream, temperature_indices, cached_energies, out);
}

int mc_chain_run(void *chain, void *stream, long long nsweeps, long long sweep0, int exchange_interval,
                 int parity, long long rounds0, int history_stride, const int32_t *tracked, int ntracked,
                 int32_t *history, long long history_capacity, long long history_row0, int *parity_out,
                 long long *rounds_out, long long *history_rows_out) {
    Chain *c = static_cast<Chain *>(chain);
    MC_REQUIRE(c, "null chain");
    MC_REQUIRE(nsweeps >= 0 && sweep0 >= 0 && rounds0 >= 0, "sweep and round counts must be nonnegative");
    MC_REQUIRE(exchange_interval >= 1, "exchange_interval must be a positive integer");
    MC_REQUIRE(c->energies, "attach the exchange buffers first (mc_chain_attach_exchange)");
    MC_REQUIRE(parity == 0 || parity == 1, "parity must be 0 or 1");
    MC_REQUIRE(!tracked || (history && history_stride >= 1 && ntracked >= 1 && history_row0 >= 0),
               "walker history needs a buffer, a positive stride and walkers");
    long long rounds = rounds0, row = history_row0;
    for (long long n = 0; n < nsweeps; n++) {
        {
            const int rc = chain_sweep(*c, stream);
            if (rc) return rc;
        }
        const long long sweep = sweep0 + n + 1;
        if (sweep % exchange_interval) continue;
        int rc = chain_energy(*c, stream, c->energies);
        if (rc) return rc;
        rc = mc_pt_exchange(stream, chain_beta_code(*c), c->energies, c->betas, c->w2t, c->t2w, c->pt_rng,
                            c->pt_counts, c->trips, c->R, parity, sweep);
        if (rc) return rc;
        rounds++;
        if (c->R > 2) parity = 1 - parity;
        if (tracked && rounds % history_stride == 0) {
            MC_REQUIRE(row < history_capacity, "walker history buffer of %lld rows is full", history_capacity);
            rc = mc_pt_record_walkers(stream, c->w2t, tracked, ntracked, history + (size_t)row * ntracked);
            if (rc) return rc;
            row++;
        }
    }
    if (parity_out) *parity_out = parity;
    if (rounds_out) *rounds_out = rounds;
    if (history_rows_out) *history_rows_out = row;
    return MC_OK;
}

}  // extern "C"
