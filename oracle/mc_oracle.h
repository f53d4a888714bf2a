/*
 * CPU oracle for the checkerboard Metropolis + parallel-tempering engine.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (JinhaoR/GPU-monte-carlo, Numba-CUDA kernels) used as the checker
 * by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The
 * product path (gpu-monte-carlo_amd/) never links, loads or calls it.
 *
 * Layouts are the reference's: state (R, L, L, ...) C-order, per-site RNG
 * streams indexed r*L*L + i*L + j with records {u64 s0, u64 s1}, measurement
 * output (C, R) float64.
 *
 * Transcendental policy (the reference's float32 kernels call NVIDIA libdevice,
 * whose rounding cannot be reproduced): a float32 transcendental (cos, sin,
 * exp, log) is one of the float32 functions of mc_oracle_fast.c (Cody-Waite
 * reduction and polynomials with explicit fmaf, at most 1.5 ulp on the
 * updates' arguments), exactly as the HIP kernels evaluate it; the Ising
 * acceptance thresholds are the float64 exp rounded once to float32.
 * float64 paths call the float64 libm function.  sqrt and division are IEEE.
 * Arithmetic is evaluated in source order without contraction
 * (-ffp-contract=off), like the reference's Python expressions.
 */
#ifndef MC_ORACLE_H
#define MC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG: numba.cuda.random xoroshiro128+ (numba 0.68.0 / 0.54.1 source) ---- */
uint64_t orc_splitmix64(uint64_t seed);
uint64_t orc_rng_next(uint64_t *state);
void orc_rng_jump(uint64_t *state);
float orc_uniform_f32(uint64_t *state);
double orc_uniform_f64(uint64_t *state);
void orc_rng_init(uint64_t *states, int64_t n, uint64_t seed, uint64_t subsequence_start);
void orc_rng_draw_u64(uint64_t *states, int64_t index, int64_t count, uint64_t *out);
void orc_rng_draw_f32(uint64_t *states, int64_t index, int64_t count, float *out);
void orc_rng_draw_f64(uint64_t *states, int64_t index, int64_t count, double *out);

void orc_set_threads(int n);

/* ---- Ising (int8 spins) ---- */
void orc_ising_color(int8_t *spins, uint64_t *rng, const float *betas,
                     int R, int L, int color, float J);
void orc_ising_sweeps(int8_t *spins, uint64_t *rng, const float *betas,
                      int R, int L, float J, int64_t nsweeps);
void orc_ising_measure(const int8_t *spins, int R, int L, float J,
                       const int32_t *tind, const double *cached, double *out);
void orc_ising_energy(const int8_t *spins, int R, int L, float J, double *out);

/* ---- XY ---- */
void orc_xy_color_f32(float *theta, uint64_t *rng, const float *betas,
                      int R, int L, int color, float J, float delta);
void orc_xy_color_f64(double *theta, uint64_t *rng, const double *betas,
                      int R, int L, int color, double J, double delta);
void orc_xy_gauss_color_f32(float *theta, uint64_t *rng, const float *betas, int R, int L, int color,
                            float J, const float *table_betas, const float *sigmas, int T,
                            int64_t *accepted);
void orc_xy_gauss_color_f64(double *theta, uint64_t *rng, const double *betas, int R, int L, int color,
                            double J, const double *table_betas, const double *sigmas, int T,
                            int64_t *accepted);
void orc_xy_sweeps_f32(float *theta, uint64_t *rng, const float *betas,
                       int R, int L, float J, float delta, int64_t nsweeps);
void orc_xy_sweeps_f64(double *theta, uint64_t *rng, const double *betas,
                       int R, int L, double J, double delta, int64_t nsweeps);
void orc_xy_measure_f32(const float *theta, int R, int L, double J,
                        const int32_t *tind, const double *cached, double *out);
void orc_xy_measure_f64(const double *theta, int R, int L, double J,
                        const int32_t *tind, const double *cached, double *out);
void orc_xy_energy_f32(const float *theta, int R, int L, double J, double *out);
void orc_xy_energy_f64(const double *theta, int R, int L, double J, double *out);

/* ---- Kawamura (5 reals per site: theta, tau, dx, dy, dz) ---- */
float orc_kawamura_local_f32(const float *state, int L, int r, int i, int j,
                             float theta, float tau, float dx, float dy,
                             float dz, float J);
double orc_kawamura_local_f64(const double *state, int L, int r, int i, int j,
                              double theta, double tau, double dx, double dy,
                              double dz, double J);
void orc_kawamura_color_f32(float *state, uint64_t *rng, const float *betas,
                            int R, int L, int color, float J, float delta);
void orc_kawamura_color_f64(double *state, uint64_t *rng, const double *betas,
                            int R, int L, int color, double J, double delta);
void orc_kawamura_sweeps_f32(float *state, uint64_t *rng, const float *betas,
                             int R, int L, float J, float delta, int64_t nsweeps);
void orc_kawamura_sweeps_f64(double *state, uint64_t *rng, const double *betas,
                             int R, int L, double J, double delta, int64_t nsweeps);
void orc_kawamura_measure_f32(const float *state, int R, int L, double J,
                              const int32_t *tind, const double *cached, double *out);
void orc_kawamura_measure_f64(const double *state, int R, int L, double J,
                              const int32_t *tind, const double *cached, double *out);
void orc_kawamura_energy_f32(const float *state, int R, int L, double J, double *out);
void orc_kawamura_energy_f64(const double *state, int R, int L, double J, double *out);

/* ---- He3 thin film (3x2 complex per site, interleaved re/im) ----
 * coef[5] = GL betas, K, g_d. */
float orc_he3_potential_f32(const float *psi, const float *coef, float g_d);
double orc_he3_potential_f64(const double *psi, const double *coef, double g_d);
float orc_he3_gradient_difference_f32(const float *state, int L, int r, int i, int j,
                                      const float *proposed);
double orc_he3_gradient_difference_f64(const double *state, int L, int r, int i, int j,
                                       const double *proposed);
float orc_he3_local_energy_f32(const float *state, int L, int r, int i, int j,
                               const float *psi, const float *coef, float K, float g_d);
void orc_he3_color_f32(float *state, uint64_t *rng, const float *betas, int R, int L,
                       int color, const float *coef, float K, float delta, float g_d);
void orc_he3_color_f64(double *state, uint64_t *rng, const double *betas, int R, int L,
                       int color, const double *coef, double K, double delta, double g_d);
void orc_he3_sweeps_f32(float *state, uint64_t *rng, const float *betas, int R, int L,
                        const float *coef, float K, float delta, float g_d, int64_t nsweeps);
void orc_he3_gaussian_color_f32(float *state, uint64_t *rng, const float *betas, int R, int L,
                                int color, const float *coef, float K, float rho, float mass,
                                int64_t *counts, float g_d);
void orc_he3_gaussian_color_f64(double *state, uint64_t *rng, const double *betas, int R, int L,
                                int color, const double *coef, double K, double rho, double mass,
                                int64_t *counts, double g_d);
void orc_he3_measure_f32(const float *state, int R, int L, const float *coef, float K, float g_d,
                         const int32_t *tind, const double *cached, double *out);
void orc_he3_measure_f64(const double *state, int R, int L, const double *coef, double K, double g_d,
                         const int32_t *tind, const double *cached, double *out);
void orc_he3_energy_f32(const float *state, int R, int L, const float *coef, float K, float g_d,
                        double *out);
void orc_he3_energy_f64(const double *state, int R, int L, const double *coef, double K, double g_d,
                        double *out);

/* ---- Parallel tempering exchange (core/parallel_tempering.py:35-91) ---- */
void orc_pt_exchange_f32(const double *energies, float *betas, int32_t *w2t, int32_t *t2w,
                         uint64_t *pt_rng, int64_t *counts, int64_t *trips, int R,
                         int parity, int64_t sweep);
void orc_pt_exchange_f64(const double *energies, double *betas, int32_t *w2t, int32_t *t2w,
                         uint64_t *pt_rng, int64_t *counts, int64_t *trips, int R,
                         int parity, int64_t sweep);

/* ---- batched float32 fast path (mc_oracle_fast.c) ---- */
void orc_fast_sincosf(float x, float *sn, float *cs);
float orc_fast_expf(float x);
void orc_fast_sincosf_n(const float *x, int64_t n, float *sn, float *cs);
void orc_fast_expf_n(const float *x, int64_t n, float *y);
float orc_fast_logf(float x);
float orc_accept_weight(float x);
void orc_fast_logf_n(const float *x, int64_t n, float *y);
void orc_fast_gauss_draw(uint64_t x, float *z, float *v);
void orc_normal_q_n(const uint32_t *k, int64_t cnt, float *y);
void orc_xyfast_gauss_color(float *th, float *c, float *s, uint64_t *rng, const float *betas, const float *sig,
                            int R, int L, int color, float J, int64_t *acc);
void orc_xyfast_init_cache(const float *th, int64_t n, float *c, float *s);
void orc_xyfast_color(float *th, float *c, float *s, uint64_t *rng, const float *betas, int R, int L,
                      int color, float J, float delta);
void orc_xyfast_sweeps(float *th, float *c, float *s, uint64_t *rng, const float *betas, int R, int L,
                       float J, float delta, int64_t nsweeps);
void orc_xyfast_observables(const float *c, const float *s, int R, int L, float J, double *out);

#ifdef __cplusplus
}
#endif
#endif
