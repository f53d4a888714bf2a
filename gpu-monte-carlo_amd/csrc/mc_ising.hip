// Ising model: per-colour Metropolis update and energy/magnetisation sums.
#include "mc_common.h"
#include "mc_reduce.h"

namespace mc {

// Colour c updates sites (i, j) = (2 row + c/2, 2 col + c%2); one thread per
// site of the colour, grid.y = replica.  dE = float(2 s sum_nn) J; one uniform
// per site is always drawn; flip iff dE <= 0 or u < exp(-beta dE) in float32.
__global__ __launch_bounds__(256) void ising_color_kernel(int8_t *__restrict__ spins,
                                                          uint64_t *__restrict__ rng,
                                                          const float *__restrict__ betas, int L,
                                                          int color, float J) {
    const int h = L >> 1;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= h * h) return;
    const int r = blockIdx.y;
    const int row = t / h, col = t - (t / h) * h;
    const int i = 2 * row + (color >> 1), j = 2 * col + (color & 1);
    int8_t *b = spins + (size_t)r * L * L;
    const size_t k = (size_t)r * L * L + (size_t)i * L + j;
    const int s = b[i * L + j];
    const int nb = (int)b[wrap_up(i, L) * L + j] + (int)b[wrap_dn(i, L) * L + j] +
                   (int)b[i * L + wrap_up(j, L)] + (int)b[i * L + wrap_dn(j, L)];
    const float dE = (float)(2 * s * nb) * J;
    Rng g = load_rng(rng, k);
    const float u = g.uniform_f32();
    store_rng(rng, k, g);
    if (dE <= 0.0f || u < exact_expf(-betas[r] * dE)) b[i * L + j] = (int8_t)(-s);
}

struct IsingAcc {
    const int8_t *spins;
    float J;
    int L;
    template <bool ENERGY_ONLY, int CA>
    __device__ __forceinline__ void site(int r, int i, int j, double (&v)[CA], bool cached) const {
        const int8_t *b = spins + (size_t)r * L * L;
        const int s = b[i * L + j];
        if (!cached) {
            const int down = b[wrap_dn(i, L) * L + j];
            const int right = b[i * L + wrap_dn(j, L)];
            v[0] += -(double)J * s * (down + right);
        }
        if (!ENERGY_ONLY) v[1] += s;
    }
    template <int CA>
    __device__ __forceinline__ void finalize(const double (&t)[CA], double energy, double *out, int R,
                                             int dest) const {
        write_plain<CA>(t, energy, out, R, dest);
    }
};

}  // namespace mc

using namespace mc;

extern "C" {

int mc_ising_update_color(void *stream, int8_t *spins, uint64_t *rng, const float *betas, int R,
                          int L, int color, float J) {
    MC_REQUIRE(L >= 2 && L % 2 == 0, "L must be even, got %d", L);
    MC_REQUIRE(R >= 1, "need at least one replica");
    MC_REQUIRE(color >= 0 && color < 4, "color must be in 0..3, got %d", color);
    const int h = L / 2;
    dim3 grid((unsigned)((h * h + 255) / 256), (unsigned)R);
    hipLaunchKernelGGL(ising_color_kernel, grid, dim3(256), 0, as_stream(stream), spins, rng, betas,
                       L, color, J);
    MC_CHECK_LAUNCH();
    return MC_OK;
}

int mc_ising_measure(void *stream, const int8_t *spins, int R, int L, float J,
                     const int32_t *temperature_indices, const double *cached_energies,
                     double *out) {
    return launch_reduce<2, false>(stream, IsingAcc{spins, J, L}, R, L, temperature_indices,
                                   cached_energies, out);
}

int mc_ising_energy(void *stream, const int8_t *spins, int R, int L, float J, double *out) {
    return launch_reduce<2, true>(stream, IsingAcc{spins, J, L}, R, L, nullptr, nullptr, out);
}

}  // extern "C"
