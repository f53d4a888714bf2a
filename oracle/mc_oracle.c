/*
 * CPU oracle for the checkerboard Metropolis + parallel-tempering engine.
 * TEST INFRASTRUCTURE ONLY: used as the checker by tests/, smoke() and the
 * cpu_baseline leg of bench.py; never linked or called by the product path.
 * See mc_oracle.h for layouts and the transcendental policy.
 */
#define _GNU_SOURCE
#include "mc_oracle.h"
#include "oracle_fastmath.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

/* ------------------------------------------------------------------ RNG
 * numba/cuda/random.py (numba 0.54.1 source text; the reference pins numba
 * 0.68.0 in profiling/HE3_BENCHMARK.md:4).  Algorithm: Blackman & Vigna,
 * xoroshiro128+ and splitmix64, http://xoroshiro.di.unimi.it/. */

static inline uint64_t rotl64(uint64_t x, unsigned k) { return (x << k) | (x >> (64u - k)); }

/* random.py:45-68: splitmix64 of the seed; s0 = s1 = z */
uint64_t orc_splitmix64(uint64_t seed) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* random.py:78-98 */
uint64_t orc_rng_next(uint64_t *st) {
    const uint64_t s0 = st[0];
    uint64_t s1 = st[1];
    const uint64_t result = s0 + s1;
    s1 ^= s0;
    st[0] = rotl64(s0, 55) ^ s1 ^ (s1 << 14);
    st[1] = rotl64(s1, 36);
    return result;
}

/* random.py:101-126: advance by 2**64 steps */
void orc_rng_jump(uint64_t *st) {
    static const uint64_t jump[2] = {0xbeac0467eba5facbULL, 0xd86b048b86aa9922ULL};
    uint64_t s0 = 0, s1 = 0;
    for (int i = 0; i < 2; i++)
        for (int b = 0; b < 64; b++) {
            if (jump[i] & (1ULL << b)) {
                s0 ^= st[0];
                s1 ^= st[1];
            }
            orc_rng_next(st);
        }
    st[0] = s0;
    st[1] = s1;
}

/* random.py:129-140: (x >> 11) * 2**-53, then round-to-nearest float32 */
double orc_uniform_f64(uint64_t *st) { return (double)(orc_rng_next(st) >> 11) * (1.0 / 9007199254740992.0); }
float orc_uniform_f32(uint64_t *st) { return (float)orc_uniform_f64(st); }

/* random.py:225-240: state 0 from the seed, then state k = jump(state k-1) */
void orc_rng_init(uint64_t *states, int64_t n, uint64_t seed, uint64_t subsequence_start) {
    if (n < 1) return;
    const uint64_t z = orc_splitmix64(seed);
    states[0] = z;
    states[1] = z;
    for (uint64_t s = 0; s < subsequence_start; s++) orc_rng_jump(states);
    for (int64_t i = 1; i < n; i++) {
        states[2 * i] = states[2 * i - 2];
        states[2 * i + 1] = states[2 * i - 1];
        orc_rng_jump(states + 2 * i);
    }
}

void orc_rng_draw_u64(uint64_t *states, int64_t index, int64_t count, uint64_t *out) {
    for (int64_t k = 0; k < count; k++) out[k] = orc_rng_next(states + 2 * index);
}
void orc_rng_draw_f32(uint64_t *states, int64_t index, int64_t count, float *out) {
    for (int64_t k = 0; k < count; k++) out[k] = orc_uniform_f32(states + 2 * index);
}
void orc_rng_draw_f64(uint64_t *states, int64_t index, int64_t count, double *out) {
    for (int64_t k = 0; k < count; k++) out[k] = orc_uniform_f64(states + 2 * index);
}

/* ------------------------------------------------------ transcendentals */
/* float32: the fast functions of mc_oracle_fast.c (the kernels' policy);
 * the Ising thresholds: float64 exp rounded once to float32 */
static inline float rcos_f32(float x) { float s, c; ofm_sincosf(x, &s, &c); return c; }
static inline float rsin_f32(float x) { float s, c; ofm_sincosf(x, &s, &c); return s; }
static inline float rexp_f32(float x) { return ofm_expf(x); }
static inline float rlog_f32(float x) { return ofm_logf(x); }
static inline float exact_expf_f32(float x) { return (float)exp((double)x); }
static inline float rsqrt_f32(float x) { return sqrtf(x); }
static inline float rfloor_f32(float x) { return floorf(x); }
static inline float rfabs_f32(float x) { return fabsf(x); }
static inline float draw_f32(uint64_t *s) { return orc_uniform_f32(s); }
static inline double rcos_f64(double x) { return cos(x); }
static inline double rsin_f64(double x) { return sin(x); }
static inline double rexp_f64(double x) { return exp(x); }
static inline double rlog_f64(double x) { return log(x); }
static inline double rsqrt_f64(double x) { return sqrt(x); }
static inline double rfloor_f64(double x) { return floor(x); }
static inline double rfabs_f64(double x) { return fabs(x); }
static inline double draw_f64(uint64_t *s) { return orc_uniform_f64(s); }

/* ---------------------------------------------------------------- Ising */

#define ISITE(L, r, i, j) ((size_t)(r) * (size_t)(L) * (size_t)(L) + (size_t)(i) * (size_t)(L) + (size_t)(j))

/* models/ising/update.py:7-29 */
void orc_ising_color(int8_t *spins, uint64_t *rng, const float *betas, int R, int L, int color,
                     float J) {
    const int h = L / 2;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < R; r++) {
        for (int row = 0; row < h; row++) {
            for (int col = 0; col < h; col++) {
                const int i = 2 * row + color / 2, j = 2 * col + color % 2;
                const size_t k = ISITE(L, r, i, j);
                const int s = spins[k];
                const int nb = spins[ISITE(L, r, (i + L - 1) % L, j)] + spins[ISITE(L, r, (i + 1) % L, j)]
                               + spins[ISITE(L, r, i, (j + L - 1) % L)] + spins[ISITE(L, r, i, (j + 1) % L)];
                const float dE = (float)(2 * s * nb) * J;
                const float u = orc_uniform_f32(rng + 2 * k);
                if (dE <= 0.0f || u < exact_expf_f32(-betas[r] * dE)) spins[k] = (int8_t)(-s);
            }
        }
    }
}

void orc_ising_sweeps(int8_t *spins, uint64_t *rng, const float *betas, int R, int L, float J,
                      int64_t nsweeps) {
    for (int64_t s = 0; s < nsweeps; s++)
        for (int c = 0; c < 4; c++) orc_ising_color(spins, rng, betas, R, L, c, J);
}

/* models/ising/measure.py:10-68 (energy, magnetization) */
void orc_ising_measure(const int8_t *spins, int R, int L, float J, const int32_t *tind,
                       const double *cached, double *out) {
    for (int r = 0; r < R; r++) {
        double e = 0, m = 0;
        for (int i = 0; i < L; i++)
            for (int j = 0; j < L; j++) {
                const int s = spins[ISITE(L, r, i, j)];
                if (!cached) {
                    const int down = spins[ISITE(L, r, (i + 1) % L, j)];
                    const int right = spins[ISITE(L, r, i, (j + 1) % L)];
                    e += -(double)J * s * (down + right);
                }
                m += s;
            }
        const int d = tind ? tind[r] : r;
        out[d] = cached ? cached[r] : e;
        out[R + d] = m;
    }
}

/* models/ising/measure.py:71-99 */
void orc_ising_energy(const int8_t *spins, int R, int L, float J, double *out) {
    for (int r = 0; r < R; r++) {
        double e = 0;
        for (int i = 0; i < L; i++)
            for (int j = 0; j < L; j++) {
                const int s = spins[ISITE(L, r, i, j)];
                e += -(double)J * s * (spins[ISITE(L, r, (i + 1) % L, j)] + spins[ISITE(L, r, i, (j + 1) % L)]);
            }
        out[r] = e;
    }
}

/* ------------------------------------------------ precision-generic part */
#define REAL float
#define SUF _f32
#include "oracle_models.h"
#undef REAL
#undef SUF

#define REAL double
#define SUF _f64
#include "oracle_models.h"
#undef REAL
#undef SUF
