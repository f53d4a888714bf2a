// XY model: per-colour uniform-window Metropolis update (float32 or float64
// state) and the five measurement sums.
#include "mc_common.h"
#include "mc_reduce.h"

namespace mc {

template <class R>
__global__ __launch_bounds__(256) void xy_color_kernel(R *__restrict__ th, uint64_t *__restrict__ rng,
                                                       const R *__restrict__ betas, int L, int color,
                                                       R J, R delta) {
    const int h = L >> 1;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= h * h) return;
    const int r = blockIdx.y;
    const int row = t / h, col = t - (t / h) * h;
    const int i = 2 * row + (color >> 1), j = 2 * col + (color & 1);
    R *b = th + (size_t)r * L * L;
    const size_t k = (size_t)r * L * L + (size_t)i * L + j;
    const R one = R(1), two = R(2), zero = R(0);
    const R pi = RealConst<R>::pi(), two_pi = RealConst<R>::two_pi();
    Rng g = load_rng(rng, k);
    const R theta = b[i * L + j];
    const R u = g.template uniform<R>();
    R p = theta + delta * (two * u - one);
    if (p >= pi)
        p -= two_pi;
    else if (p < -pi)
        p += two_pi;
    const R up = b[wrap_up(i, L) * L + j];
    const R down = b[wrap_dn(i, L) * L + j];
    const R left = b[i * L + wrap_up(j, L)];
    const R right = b[i * L + wrap_dn(j, L)];
    const R e_old = -J * (((rcos(theta - up) + rcos(theta - down)) + rcos(theta - left)) + rcos(theta - right));
    const R e_new = -J * (((rcos(p - up) + rcos(p - down)) + rcos(p - left)) + rcos(p - right));
    const R dE = e_new - e_old;
    const R v = g.template uniform<R>();
    store_rng(rng, k, g);
    if (metropolis_accept(dE, v, betas[r])) b[i * L + j] = p;
}

// Angle into [-pi, pi) after an unbounded move: subtract whole turns, then
// fix the rounding of the last one.
template <class R>
__device__ __forceinline__ R wrap_angle(R p, R pi, R two_pi) {
    if (p >= pi || p < -pi) {
        p -= two_pi * ::floor((p + pi) / two_pi);
        if (p >= pi)
            p -= two_pi;
        else if (p < -pi)
            p += two_pi;
    }
    return p;
}

// Index of the ladder temperature whose beta is nearest to `beta` (first on ties).
template <class R>
__device__ __forceinline__ int nearest_temperature(const R *table, int T, R beta) {
    int best = 0;
    R bd = ::fabs(table[0] - beta);
    for (int t = 1; t < T; t++) {
        const R d = ::fabs(table[t] - beta);
        if (d < bd) {
            bd = d;
            best = t;
        }
    }
    return best;
}

// Gaussian-proposal XY update (no reference counterpart; the reference draws
// a uniform window, models/xy/update.py:24-31).  Per site: u1, u2 ->
// z = sqrt(-2 log u1) cos(2 pi u2) (u1 = 0 replaced by 2^-53), p = theta +
// sigma_t z wrapped to [-pi, pi), then the reference's dE and acceptance with
// a third draw.  sigma_t belongs to the ladder temperature t whose beta is
// nearest to the walker's, so sigmas follow temperatures through exchanges;
// accepted moves are counted per temperature.
template <class R>
__global__ __launch_bounds__(256) void xy_gauss_kernel(R *__restrict__ th, uint64_t *__restrict__ rng,
                                                       const R *__restrict__ betas, int L, int color, R J,
                                                       const R *__restrict__ table_betas,
                                                       const R *__restrict__ sigmas, int T,
                                                       unsigned long long *__restrict__ accepted) {
    __shared__ int t_idx;
    const int r = blockIdx.y;
    if (threadIdx.x == 0) t_idx = nearest_temperature(table_betas, T, betas[r]);
    __syncthreads();
    const int h = L >> 1;
    const int t = blockIdx.x * 256 + threadIdx.x;
    int acc = 0;
    if (t < h * h) {
        const int row = t / h, col = t - (t / h) * h;
        const int i = 2 * row + (color >> 1), j = 2 * col + (color & 1);
        R *b = th + (size_t)r * L * L;
        const size_t k = (size_t)r * L * L + (size_t)i * L + j;
        const R two = R(2), zero = R(0);
        const R pi = RealConst<R>::pi(), two_pi = RealConst<R>::two_pi();
        Rng g = load_rng(rng, k);
        const R theta = b[i * L + j];
        R u1 = g.template uniform<R>();
        const R u2 = g.template uniform<R>();
        if (u1 <= zero) u1 = R(0x1p-53);
        const R z = rsqrt_(-two * rlog(u1)) * rcos(two_pi * u2);
        const R p = wrap_angle(theta + sigmas[t_idx] * z, pi, two_pi);
        const R up = b[wrap_up(i, L) * L + j];
        const R down = b[wrap_dn(i, L) * L + j];
        const R left = b[i * L + wrap_up(j, L)];
        const R right = b[i * L + wrap_dn(j, L)];
        const R e_old = -J * (((rcos(theta - up) + rcos(theta - down)) + rcos(theta - left)) + rcos(theta - right));
        const R e_new = -J * (((rcos(p - up) + rcos(p - down)) + rcos(p - left)) + rcos(p - right));
        const R dE = e_new - e_old;
        const R v = g.template uniform<R>();
        store_rng(rng, k, g);
        if (metropolis_accept(dE, v, betas[r])) {
            b[i * L + j] = p;
            acc = 1;
        }
    }
    if (accepted) {
        const int n = __syncthreads_count(acc);
        if (threadIdx.x == 0 && n) atomicAdd(&accepted[t_idx], (unsigned long long)n);
    }
}

template <class R>
static int xy_gauss_update(void *stream, void *theta, uint64_t *rng, const void *betas, int R_, int L, int color,
                           double J, const void *table_betas, const void *sigmas, int T,
                           unsigned long long *accepted) {
    const int h = L / 2;
    dim3 grid((unsigned)((h * h + 255) / 256), (unsigned)R_);
    hipLaunchKernelGGL(xy_gauss_kernel<R>, grid, dim3(256), 0, as_stream(stream), (R *)theta, rng,
                       (const R *)betas, L, color, (R)J, (const R *)table_betas, (const R *)sigmas, T, accepted);
    MC_CHECK_LAUNCH();
    return MC_OK;
}

template <class R>
struct XYAcc {
    const R *th;
    double J;
    int L;
    template <bool ENERGY_ONLY, int CA>
    __device__ __forceinline__ void site(int r, int i, int j, double (&v)[CA], bool cached) const {
        const R *b = th + (size_t)r * L * L;
        const double theta = (double)b[i * L + j];
        const double tx = (double)b[wrap_dn(i, L) * L + j];
        const double dx = theta - tx;
        const double cos_x = ::cos(dx);
        if (!cached) {
            const double ty = (double)b[i * L + wrap_dn(j, L)];
            v[0] += -J * (cos_x + ::cos(theta - ty));
        }
        if (!ENERGY_ONLY) {
            v[1] += ::cos(theta);
            v[2] += ::sin(theta);
            v[3] += cos_x;
            v[4] += ::sin(dx);
        }
    }
    template <int CA>
    __device__ __forceinline__ void finalize(const double (&t)[CA], double energy, double *out, int R_,
                                             int dest) const {
        write_plain<CA>(t, energy, out, R_, dest);
    }
};

template <class R>
static int xy_update(void *stream, void *theta, uint64_t *rng, const void *betas, int R_, int L,
                     int color, double J, double delta) {
    const int h = L / 2;
    dim3 grid((unsigned)((h * h + 255) / 256), (unsigned)R_);
    hipLaunchKernelGGL(xy_color_kernel<R>, grid, dim3(256), 0, as_stream(stream), (R *)theta, rng,
                       (const R *)betas, L, color, (R)J, (R)delta);
    MC_CHECK_LAUNCH();
    return MC_OK;
}

}  // namespace mc

using namespace mc;

extern "C" {

int mc_xy_update_color(void *stream, int dtype, void *theta, uint64_t *rng, const void *betas, int R,
                       int L, int color, double J, double delta) {
    MC_REQUIRE(L >= 2 && L % 2 == 0, "L must be even, got %d", L);
    MC_REQUIRE(R >= 1, "need at least one replica");
    MC_REQUIRE(color >= 0 && color < 4, "color must be in 0..3, got %d", color);
    if (dtype == MC_F32) return xy_update<float>(stream, theta, rng, betas, R, L, color, J, delta);
    if (dtype == MC_F64) return xy_update<double>(stream, theta, rng, betas, R, L, color, J, delta);
    set_error(MC_ERR_TYPE, "unsupported XY dtype code %d", dtype);
    return MC_ERR_TYPE;
}

int mc_xy_gauss_update_color(void *stream, int dtype, void *theta, uint64_t *rng, const void *betas, int R,
                             int L, int color, double J, const void *table_betas, const void *sigmas, int T,
                             unsigned long long *accepted) {
    MC_REQUIRE(L >= 2 && L % 2 == 0, "L must be even, got %d", L);
    MC_REQUIRE(R >= 1 && T >= 1, "need at least one replica and one temperature");
    MC_REQUIRE(color >= 0 && color < 4, "color must be in 0..3, got %d", color);
    MC_REQUIRE(table_betas && sigmas, "the temperature table and sigmas are required");
    if (dtype == MC_F32)
        return xy_gauss_update<float>(stream, theta, rng, betas, R, L, color, J, table_betas, sigmas, T, accepted);
    if (dtype == MC_F64)
        return xy_gauss_update<double>(stream, theta, rng, betas, R, L, color, J, table_betas, sigmas, T, accepted);
    set_error(MC_ERR_TYPE, "unsupported XY dtype code %d", dtype);
    return MC_ERR_TYPE;
}

int mc_xy_measure(void *stream, int dtype, const void *theta, int R, int L, double J,
                  const int32_t *temperature_indices, const double *cached_energies, double *out) {
    if (dtype == MC_F32)
        return launch_reduce<5, false>(stream, XYAcc<float>{(const float *)theta, J, L}, R, L,
                                       temperature_indices, cached_energies, out);
    if (dtype == MC_F64)
        return launch_reduce<5, false>(stream, XYAcc<double>{(const double *)theta, J, L}, R, L,
                                       temperature_indices, cached_energies, out);
    set_error(MC_ERR_TYPE, "unsupported XY dtype code %d", dtype);
    return MC_ERR_TYPE;
}

int mc_xy_energy(void *stream, int dtype, const void *theta, int R, int L, double J, double *out) {
    if (dtype == MC_F32)
        return launch_reduce<5, true>(stream, XYAcc<float>{(const float *)theta, J, L}, R, L, nullptr,
                                      nullptr, out);
    if (dtype == MC_F64)
        return launch_reduce<5, true>(stream, XYAcc<double>{(const double *)theta, J, L}, R, L, nullptr,
                                      nullptr, out);
    set_error(MC_ERR_TYPE, "unsupported XY dtype code %d", dtype);
    return MC_ERR_TYPE;
}

}  // extern "C"
