/*
 * mcgpu — C ABI of the gfx950 (MI355X) checkerboard Metropolis +
 * parallel-tempering engine.
 *
 * Plain pointers and sizes only.  Every device pointer is memory obtained
 * from mc_malloc (or any hipMalloc'd buffer); `stream` is a hipStream_t
 * (NULL = the default stream).  Every call returns MC_OK or an error code;
 * mc_last_error() gives the message.  Launches are asynchronous on `stream`;
 * the mc_memcpy_* calls are synchronous.  Calls on one stream execute in
 * order; the library is not thread-safe (callers serialise).
 *
 * Layouts follow the reference (JinhaoR/GPU-monte-carlo):
 *   state      (R, L, L[, comps]) C-order; Ising int8, XY real, Kawamura
 *              5 reals (theta, tau, dx, dy, dz), He3 3x2 complex (re, im)
 *   rng        R*L*L records {uint64 s0, s1}, stream k = r*L*L + i*L + j
 *   out        measurement channels (C, R) float64
 *   dtype      MC_F32 (reference precision) or MC_F64 (float64 mode)
 * The reference binds none of this through an FFI; its kernels are Numba
 * launches.  Each entry point names the reference function it replaces.
 */
#ifndef MCGPU_H
#define MCGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MC_OK = 0,
    MC_ERR_VALUE = 1,      /* invalid argument  -> ValueError */
    MC_ERR_HIP = 2,        /* HIP runtime error -> RuntimeError */
    MC_ERR_NOTIMPL = 3,    /* missing path      -> NotImplementedError */
    MC_ERR_TYPE = 4        /* wrong model/dtype -> TypeError */
};

enum { MC_F32 = 0, MC_F64 = 1 };

/* ---------------------------------------------------------------- runtime */

/* Message of the last failed call on this thread. */
const char *mc_last_error(void);
/* Library version string. */
const char *mc_version(void);
/* Number of visible HIP devices (0 when none). */
int mc_device_count(int *count);
/* Make `device` current and prepare per-device tables. */
int mc_init(int device);

/* Device memory: replaces numba cuda.device_array / cuda.to_device /
 * DeviceNDArray.copy_to_host / copy_to_device (core/runtime.py:109-119,
 * core/parallel_tempering.py:151,171-188). */
int mc_malloc(void **ptr, size_t bytes);
int mc_free(void *ptr);
/* Pinned (page-locked) host memory, zeroed, device-accessible: targets of
 * asynchronous copies and of device stores (mc_batch_run's error_host). */
int mc_host_alloc(void **ptr, size_t bytes);
int mc_host_free(void *ptr);
int mc_memcpy_htod(void *dst, const void *src, size_t bytes);
int mc_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int mc_memcpy_dtod(void *dst, const void *src, size_t bytes);
int mc_memset(void *dst, int value, size_t bytes);
int mc_synchronize(void);
int mc_stream_create(void **stream);
int mc_stream_destroy(void *stream);
int mc_stream_synchronize(void *stream);
/* HIP events for timing work on a given stream (no reference counterpart:
 * the reference times with cuda.synchronize + perf_counter,
 * profiling/profile_he3.py:112-133). */
int mc_event_create(void **event);
int mc_event_destroy(void *event);
int mc_event_record(void *event, void *stream);
int mc_event_elapsed_ms(void *start, void *stop, float *ms);
/* Name and compute-unit count of the current device. */
int mc_device_info(char *name, int name_bytes, int *compute_units);

/* ------------------------------------------------------------------- RNG */

/* numba create_xoroshiro128p_states(n, seed, subsequence_start)
 * (core/runtime.py:113-115, core/parallel_tempering.py:177-179;
 * numba/cuda/random.py:225-291): state k = jump^(k + subsequence_start)
 * of splitmix64(seed).  Computed on the GPU in parallel. */
int mc_rng_init(void *stream, uint64_t *states, int64_t n, uint64_t seed,
                uint64_t subsequence_start);

/* --------------------------------------------------------------- Ising */

/* models/ising/update.py:7-29 (metropolis_color), models/ising/model.py:19-22 */
int mc_ising_update_color(void *stream, int8_t *spins, uint64_t *rng, const float *betas,
                          int R, int L, int color, float J);
/* models/ising/measure.py:10-68 (energy, magnetization); temperature_indices
 * and cached_energies may be NULL (the four kernel variants). */
int mc_ising_measure(void *stream, const int8_t *spins, int R, int L, float J,
                     const int32_t *temperature_indices, const double *cached_energies,
                     double *out);
/* models/ising/measure.py:71-99 (ising_energy_kernel) */
int mc_ising_energy(void *stream, const int8_t *spins, int R, int L, float J, double *out);

/* ------------------------------------------------------------------ XY */

/* models/xy/update.py:6-54; betas have the state's dtype */
int mc_xy_update_color(void *stream, int dtype, void *theta, uint64_t *rng, const void *betas,
                       int R, int L, int color, double J, double delta);
/* Gaussian-proposal XY update; no reference counterpart (the reference's
 * uniform window is models/xy/update.py:24-31, replaced here by
 * theta + sigma_t z, z = sqrt(-2 log u1) cos(2 pi u2), wrapped by whole turns;
 * dE and acceptance as models/xy/update.py:33-54 with a third draw).
 * table_betas and sigmas (T entries, state dtype) give each ladder
 * temperature's sigma; a walker uses the temperature whose beta is nearest
 * to its own.  accepted (T entries, may be NULL) counts accepted moves per
 * temperature. */
int mc_xy_gauss_update_color(void *stream, int dtype, void *theta, uint64_t *rng, const void *betas,
                             int R, int L, int color, double J, const void *table_betas,
                             const void *sigmas, int T, unsigned long long *accepted);
/* models/xy/measure.py:17-68 (energy, mx, my, twist_cos_x, twist_sin_x) */
int mc_xy_measure(void *stream, int dtype, const void *theta, int R, int L, double J,
                  const int32_t *temperature_indices, const double *cached_energies,
                  double *out);
/* models/xy/measure.py:71-96 */
int mc_xy_energy(void *stream, int dtype, const void *theta, int R, int L, double J,
                 double *out);

/* ------------------------------------------------------------ Kawamura */

/* models/kawamura/update.py:21-65 */
int mc_kawamura_update_color(void *stream, int dtype, void *state, uint64_t *rng,
                             const void *betas, int R, int L, int color, double J,
                             double delta);
/* One sweep: colours 0..3 of models/kawamura/update.py:21-65 in order, as two
 * launches (colours 0+1 of the even rows, then 2+3 of the odd rows) that
 * stage whole rows in LDS; bit-identical to four mc_kawamura_update_color
 * calls. */
int mc_kawamura_sweep(void *stream, int dtype, void *state, uint64_t *rng, const void *betas, int R,
                      int L, double J, double delta);
/* models/kawamura/measure.py:14-70 (energy, chirality, px, py, spin2,
 * twist_cos_x, twist_sin_x) */
int mc_kawamura_measure(void *stream, int dtype, const void *state, int R, int L, double J,
                        const int32_t *temperature_indices, const double *cached_energies,
                        double *out);
/* models/kawamura/measure.py:71-99 */
int mc_kawamura_energy(void *stream, int dtype, const void *state, int R, int L, double J,
                       double *out);

/* ----------------------------------------------------------------- He3
 * coef = the five GL betas (host array, state precision as double). */

/* models/he3/update.py:176-227 (Cartesian Metropolis) */
int mc_he3_update_color(void *stream, int dtype, void *state, uint64_t *rng, const void *betas,
                        int R, int L, int color, const double *coef, double K, double delta,
                        double g_d);
/* One sweep of the Cartesian update: core/runtime.py:87-99 (colours 0..3 of
 * models/he3/update.py:176-227) in two launches (colours 0+1 of the even rows,
 * then 2+3 of the odd rows, a workgroup owning whole rows); bit-identical to
 * four mc_he3_update_color calls. */
int mc_he3_sweep(void *stream, int dtype, void *state, uint64_t *rng, const void *betas, int R, int L,
                 const double *coef, double K, double delta, double g_d);
/* models/he3/gaussian_fast.py:91-119 (conditional-Gaussian MH); counts is
 * int64 (3, R, L, L) */
int mc_he3_gaussian_update_color(void *stream, int dtype, void *state, uint64_t *rng,
                                 const void *betas, int R, int L, int color,
                                 const double *coef, double K, double rho, double mass,
                                 int64_t *counts, double g_d);
/* models/he3/measure.py:22-89 (10 channels) */
int mc_he3_measure(void *stream, int dtype, const void *state, int R, int L,
                   const double *coef, double K, double g_d,
                   const int32_t *temperature_indices, const double *cached_energies,
                   double *out);
/* models/he3/measure.py:115-143 */
int mc_he3_energy(void *stream, int dtype, const void *state, int R, int L,
                  const double *coef, double K, double g_d, double *out);

/* ------------------------------------------------ parallel tempering */

/* core/parallel_tempering.py:35-91 (_exchange_kernel): one matching of
 * neighbour pairs of `parity`, then round-trip bookkeeping.  betas has
 * `beta_dtype`; counts int64 (3, R-1); trips int64 (5, R). */
int mc_pt_exchange(void *stream, int beta_dtype, const double *energies, void *betas,
                   int32_t *walker_to_temperature, int32_t *temperature_to_walker,
                   uint64_t *pt_rng, int64_t *counts, int64_t *trips, int R, int parity,
                   int64_t sweep);
/* core/parallel_tempering.py:28-32 (_record_walkers_kernel) */
int mc_pt_record_walkers(void *stream, const int32_t *walker_to_temperature,
                         const int32_t *walker_ids, int n, int32_t *history_row);

/* ------------------------------------------------------- native chains
 * The drop-in loop of core/runtime.py:87-99 (Runtime.sweep: colours 0..3 of
 * model.update_color) and core/parallel_tempering.py:245-269 (sweep, then
 * measure_energy + _exchange_kernel every exchange_interval sweeps, and
 * _record_walkers_kernel every history_stride rounds) for the built-in
 * models, run from C: one call queues any number of sweeps.  A chain binds
 * device buffers owned by the caller (state, rng, betas; the PT arrays with
 * mc_chain_attach_exchange) and allocates nothing.
 * kind / params / buffers:
 *   0 Ising            J                                   -
 *   1 XY               J, delta                            -
 *   2 XY Gaussian      J, T                                table_betas, sigmas, accepted
 *   3 Kawamura         J, delta                            -
 *   4 He3 Cartesian    c0..c4, K, delta, g_d               -
 *   5 He3 Gaussian     c0..c4, K, rho, mass, g_d           counts (int64 (3, R, L, L))
 * dtype is the state precision (MC_F32 for Ising). */
int mc_chain_create(int kind, int dtype, int L, int R, void *state, uint64_t *rng, void *betas,
                    const double *params, int nparams, void *const *buffers, int nbuffers, void **chain);
int mc_chain_destroy(void *chain);
/* Measurement channels of the chain's model (Ising 2, XY 5, Kawamura 7, He3 10) */
int mc_chain_channels(void *chain, int *channels);
int mc_chain_attach_exchange(void *chain, double *energies, int32_t *walker_to_temperature,
                             int32_t *temperature_to_walker, uint64_t *pt_rng, int64_t *counts,
                             int64_t *trips);
/* nsweeps sweeps: colours 0, 1, 2, 3 each */
int mc_chain_sweeps(void *chain, void *stream, long long nsweeps);
/* The model's measure_energy / measure (as the mc_*_energy, mc_*_measure
 * entry points) into device buffers */
int mc_chain_energy(void *chain, void *stream, double *out);
/* mc_chain_energy that allocates nothing.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Its reduction partials live in a static
 * device array of 65,536 float64 per GPU (512 KiB, held until the process
 * ends): R x (blocks per replica: ceil(L*L / 2048), at most 64) must not
 * exceed 65,536, else MC_ERR_VALUE before any launch.  Calls on one GPU share
 * the array, so calls that overlap on one GPU (on different streams) race. */
int mc_chain_energy_noalloc(void *chain, void *stream, double *out);
int mc_chain_measure(void *chain, void *stream, const int32_t *temperature_indices,
                     const double *cached_energies, double *out);
/* nsweeps sweeps from global sweep sweep0; after every sweep s with
 * s % exchange_interval == 0: energies, one matching of `parity` (flipped
 * after each round when R > 2), and, when tracked is not NULL, the tracked
 * walkers' temperatures into history row history_row0 + k every
 * history_stride rounds (rows of ntracked int32; at most history_capacity).
 * Returns the next parity, the round count and the next history row. */
int mc_chain_run(void *chain, void *stream, long long nsweeps, long long sweep0, int exchange_interval,
                 int parity, long long rounds0, int history_stride, const int32_t *tracked, int ntracked,
                 int32_t *history, long long history_capacity, long long history_row0, int *parity_out,
                 long long *rounds_out, long long *history_rows_out);

/* -------------------------------------------- differentiable energies
 * No reference counterpart: the reference measures the twist sums only as
 * observables (models/xy/measure.py:17-68, models/kawamura/measure.py:14-70,
 * used by models/xy/analysis.py:47-54 and models/kawamura/analysis.py:36-53).
 * With a twist Delta_r on the x bonds (axis 1) and bond weights w (1 for XY,
 * (1 + tau tau')(d . d') for Kawamura), a = theta - theta_neighbour:
 *   Cx = sum_x w cos(a - Delta), Cy = sum_y w cos(a), Sx = sum_x w sin(a - Delta),
 *   H = -J (Cx + Cy).
 * twist may be NULL (all zero); every coefficient array may be NULL (zero).
 * terms: out is (3, R) float64 = (Cx, Cy, Sx).
 * grad: per-site state gradient of alpha_r Cx + beta_r Cy + gamma_r Sx, same
 * layout and dtype as the state.
 * hvp (XY): per-site Hessian of that combination times u. */
int mc_xy_twist_terms(void *stream, int dtype, const void *theta, int R, int L, const double *twist,
                      double *out);
int mc_xy_twist_grad(void *stream, int dtype, const void *theta, int R, int L, const double *twist,
                     const double *alpha, const double *beta, const double *gamma, void *grad);
int mc_xy_twist_hvp(void *stream, int dtype, const void *theta, int R, int L, const double *twist,
                    const double *alpha, const double *beta, const double *gamma, const void *u,
                    void *out);
int mc_kawamura_twist_terms(void *stream, int dtype, const void *state, int R, int L,
                            const double *twist, double *out);
int mc_kawamura_twist_grad(void *stream, int dtype, const void *state, int R, int L,
                           const double *twist, const double *alpha, const double *beta,
                           const double *gamma, void *grad);

/* ---------------------------------------------- batched persistent path
 * No single reference function: this runs core/simulation.py:36-72's loop
 * (sweeps from core/runtime.py:87-99, exchanges from
 * core/parallel_tempering.py:245-269, samples every sample_interval) for
 * one lattice size inside one persistent launch.  model 0 = Ising (the
 * reference update, bit-identical to mc_ising_update_color), model 1 = XY
 * float32 fast update (cached cos/sin, float32 sincos/exp), model 2 = the
 * same with Gaussian proposals (sigmas: R entries per ladder temperature; z
 * by inverse-CDF sampling from one 64-bit draw per site).
 * accepted (R entries per ladder temperature, may be NULL) counts accepted
 * moves.  The state is split into L/H strips of H rows per replica; each
 * workgroup (threads = 256 or 512) keeps its strip on chip.  Workspace sizes
 * come from mc_batch_workspace_size.  With exchanges, betas must hold the
 * ladder betas in walker order (betas[w] = ladder[walker_to_temperature[w]]),
 * as every exchange leaves them; counts are added to in place. */
int mc_batch_workspace_size(int model, int L, int R, int H, size_t *halo_words,
                            size_t *energy_words);
/* Largest grid (R * L / H workgroups) that can be co-resident for this
 * model and strip height. */
int mc_batch_max_workgroups(int model, int L, int R, int H, int threads, int exchange,
                            int *max_workgroups);
int mc_batch_run(void *stream, int model, int L, int R, int H, int threads, void *state, uint64_t *rng,
                 float *betas, float J, float delta, const float *sigmas,
                 unsigned long long *accepted, int nsweeps, long long sweep0,
                 int exchange_interval, int *parity, double *energies, int32_t *walker_to_temperature,
                 int32_t *temperature_to_walker, uint64_t *pt_rng, long long *counts,
                 long long *trips, int sample_interval, long long sample_first, int max_samples,
                 double *sample_partials, int32_t *sample_w2t, uint64_t *halo, uint64_t *energy,
                 int *error, long long *clock, int *error_host);
/* halo, energy and error form the launch's workspace (error: one int, the
 * watchdog code, nonzero when a bounded wait gave up).  The energy words end
 * with a staging area for the PT state at the end of the launch (betas, maps,
 * PT streams, parity); with exchanges, a small kernel queued after the launch
 * copies it into betas, walker_to_temperature, temperature_to_walker, pt_rng
 * and parity, so that no workgroup can read the state of the launch's end
 * (every workgroup reads the shared PT state when it starts).
 * clock (host memory, one long long, or NULL) counts the tags issued into the
 * workspace since it was zeroed, so that launches need not clear it: zero it
 * with a new workspace, and again after an error; a launch zeroes the
 * workspace itself (one small kernel first) when clock is NULL or reads 0,
 * and before the count could overflow (at most 2^30 - 2 sweeps per launch),
 * and advances it.  One clock per workspace; launches on one workspace run in
 * order (one stream).
 * error_host (pinned host memory from mc_host_alloc, or NULL) receives
 * error[0] from the small kernel queued on the stream after the launch (a
 * device store into the pinned word): read it after synchronising the
 * stream. */
/* The same launch in two calls, no single counterpart in the reference (the
 * loop of core/simulation.py:36-72 calls its runtime once per segment):
 * mc_batch_plan_create checks the arguments of mc_batch_run that stay the
 * same from one launch of a run to the next and resolves the kernel, its LDS
 * and the co-residency check once; mc_batch_plan_launch then queues
 * nsweeps sweeps exactly as mc_batch_run with the same arguments would, from
 * eight arguments.  The plan keeps the pointers it was given (the caller's
 * arrays, workspace, clock and error_host must outlive it); J and delta are
 * fixed at creation.  A plan is not thread-safe; destroy it with
 * mc_batch_plan_destroy. */
int mc_batch_plan_create(void *stream, int model, int L, int R, int H, int threads, void *state,
                         uint64_t *rng, float *betas, float J, float delta, const float *sigmas,
                         unsigned long long *accepted, int exchange_interval, int *parity,
                         double *energies, int32_t *walker_to_temperature,
                         int32_t *temperature_to_walker, uint64_t *pt_rng, long long *counts,
                         long long *trips, uint64_t *halo, uint64_t *energy, int *error,
                         long long *clock, int *error_host, void **plan);
int mc_batch_plan_launch(void *plan, int nsweeps, long long sweep0, int sample_interval,
                         long long sample_first, int max_samples, double *sample_partials,
                         int32_t *sample_w2t);
int mc_batch_plan_destroy(void *plan);
/* Test hook, no counterpart in the reference: later launches make each
 * strip's halo receives wait for tags that no workgroup publishes (when
 * `unreachable` is nonzero and a replica has more than one strip), and every
 * bounded wait give up after `polls` polls (0: the default, about a second).
 * (0, 0) restores normal launches.  Lets a test force the watchdog path with
 * every kernel still exiting normally. */
int mc_batch_test_watchdog(int unreachable, unsigned polls);
/* Launch mode of later mc_batch_run calls, no counterpart in the reference:
 * 0 (the default) a plain dispatch after a co-residency check (occupancy x
 * compute units) on the host, nonzero a cooperative launch. */
int mc_batch_set_cooperative(int cooperative);
/* Energies of the batched and packed paths are integer sums: Ising
 * sum s (s_down + s_right) exactly, XY the per-site (cos bond x + cos bond y)
 * in float32 rounded to multiples of 2^-40 and added in int64; the energy is
 * -J times that sum (times 2^-40 for XY).  sample_partials channel 0 holds
 * each strip's energy. */

/* ---------------------------------------------- packed persistent path
 * No single reference function: the loop of core/simulation.py:36-72 (as in
 * mc_batch_run) for several ensembles -- lattice sizes, models, ladders --
 * in one launch, e.g. a finite-size-scaling study.  A pack owns its
 * workspace; the ensembles' arrays stay the caller's.  Each ensemble is cut
 * into strips of H rows; a task is one strip position of K replicas and is
 * streamed through LDS every sweep.  Sequence: mc_pack_create, one
 * mc_pack_add_group per ensemble, mc_pack_set_work, then mc_pack_run (or
 * mc_pack_capture + mc_graph_launch) any number of times, mc_pack_destroy.
 * Trajectories, exchanges and samples equal mc_batch_run's for the same
 * ensemble. */
/* Dynamic LDS of a workgroup running tasks of shape (H, K) of an ensemble */
int mc_pack_lds_bytes(int model, int L, int R, int H, int K, size_t *bytes);
/* Co-resident packed workgroups of `threads` (256 or 512) threads with
 * lds_bytes of dynamic LDS */
int mc_pack_occupancy(int threads, size_t lds_bytes, int *max_workgroups);
int mc_pack_create(int threads, void **pack);
int mc_pack_destroy(void *pack);
/* Arguments as mc_batch_run's; K replicas per task (a power of two dividing
 * R, at most threads / 16), L <= threads.  group_index receives the
 * ensemble's index in the pack. */
int mc_pack_add_group(void *pack, int model, int L, int R, int H, int K, int exchange_interval, float J,
                      float delta, void *state, uint64_t *rng, float *betas, const float *sigmas,
                      unsigned long long *accepted, int *parity, double *energies,
                      int32_t *walker_to_temperature, int32_t *temperature_to_walker, uint64_t *pt_rng,
                      long long *counts, long long *trips, int *group_index);
/* work: nworkgroups rows of (group, first task, task count); a group's rows
 * are consecutive and list its tasks (R / K * L / H of them) in order.  The
 * launch must fit co-resident (mc_pack_occupancy). */
int mc_pack_set_work(void *pack, const int *work, int nworkgroups);
/* The sweep count the next launch starts from (default 0); it lives on the
 * device and every launch or replay advances it. */
int mc_pack_set_sweeps(void *pack, long long sweeps);
int mc_pack_sweeps(void *pack, long long *sweeps);
/* Nonzero after a launch stopped by its watchdog (synchronous) */
int mc_pack_error(void *pack, int *error);
/* nsweeps sweeps of every ensemble.  sample_partials[g] ([max_samples][R][L/H][5]
 * float64) and sample_w2t[g] ([max_samples][R]) are device buffers per
 * ensemble, in a host array; samples at sweeps sample_first + k * interval. */
int mc_pack_run(void *pack, void *stream, int nsweeps, int sample_interval, long long sample_first,
                int max_samples, double *const *sample_partials, int32_t *const *sample_w2t);
/* The same launch captured into an executable HIP graph (prologue kernel +
 * packed kernel, no memset nodes); each mc_graph_launch continues the chains
 * by nsweeps.  Samples of all replays go to the given buffers: their
 * addresses are launch arguments, fixed in the graph at capture, so later
 * runs or captures with other buffers do not change where a graph writes.
 * At most 64 groups per pack. */
int mc_pack_capture(void *pack, void *stream, int nsweeps, int sample_interval, long long sample_first,
                    int max_samples, double *const *sample_partials, int32_t *const *sample_w2t,
                    void **graph);
int mc_graph_launch(void *graph, void *stream);
int mc_graph_destroy(void *graph);

#ifdef __cplusplus
}
#endif
#endif
