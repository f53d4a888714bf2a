// Deterministic per-replica reductions for the measurement and energy kernels.
//
// Two launches.  reduce_kernel: grid (nblk, R) x 256 threads; each block
// strides over a replica's sites, reduces its float64 channels with wavefront
// shuffles and LDS and writes one partial per channel.  reduce_finish_kernel:
// one wavefront per replica adds the partials in a fixed order and writes the
// outputs.  The kernel boundary orders the partials before their sum, so no
// block needs a device-scope acquire/release (on gfx950 each of those writes
// back and invalidates the L2).  The order of every addition is fixed by
// (L, nblk), so the same state always gives bit-identical results, and an
// energy-only launch reproduces channel 0 of the full measurement exactly.
#pragma once

#include "mc_common.h"

namespace mc {

constexpr int kReduceThreads = 256;

// Blocks per replica: at least 2 sites per thread, at most 64 blocks (a
// function of L alone, so the order of the additions does not depend on the
// device).  With 8 sites per thread the He3 energy of L = 64 and 100
// replicas ran as 200 workgroups, fewer than the compute units, at 12.5 us.
inline int reduce_blocks(int L) {
    const long sites = (long)L * L;
    long b = (sites + kReduceThreads * 2 - 1) / (kReduceThreads * 2);
    if (b < 1) b = 1;
    if (b > 64) b = 64;
    return (int)b;
}

// wave_sum (lane 0's sum of a wavefront) is in mc_common.h.

// One block's C channel sums (v per thread) into partials[0..C-1] (thread 0).
template <int C, int NT>
__device__ __forceinline__ void block_partial(const double (&v)[C], double *partials) {
    __shared__ double sh[C][NT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < C; c++) {
        const double s = wave_sum(v[c]);
        if (lane == 0) sh[c][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < C; c++) {
            double t = sh[c][0];
#pragma unroll
            for (int w = 1; w < NT / 64; w++) t += sh[c][w];
            partials[c] = t;
        }
    }
}

// Acc must provide:
//   template <bool ENERGY_ONLY> __device__ void site(int r, int i, int j,
//        double (&v)[CA], bool cached) const;
//   __device__ void finalize(const double (&t)[CA], double energy,
//        double *out, int R, int dest) const;   (full measurement only)
template <int CA, bool ENERGY_ONLY, class Acc>
__global__ __launch_bounds__(kReduceThreads) void reduce_kernel(Acc acc, int L, int nblk, bool use_cache,
                                                                double *__restrict__ partials) {
    constexpr int C = ENERGY_ONLY ? 1 : CA;
    const int r = blockIdx.y, b = blockIdx.x;
    double v[CA];
#pragma unroll
    for (int c = 0; c < CA; c++) v[c] = 0.0;
    // sites k = b * 256 + tid + m * stride in order, with (i, j) = divmod(k, L)
    // advanced by the constant step divmod(stride, L) instead of a division
    // per site
    const int n = L * L, stride = nblk * kReduceThreads;
    const int di = stride / L, dj = stride - di * L;
    int k = b * kReduceThreads + threadIdx.x;
    int i = k / L, j = k - i * L;
    for (; k < n; k += stride) {
        acc.template site<ENERGY_ONLY>(r, i, j, v, use_cache);
        i += di;
        j += dj;
        if (j >= L) {
            j -= L;
            i += 1;
        }
    }
    double w[C];
#pragma unroll
    for (int c = 0; c < C; c++) w[c] = v[c];
    block_partial<C, kReduceThreads>(w, partials + ((size_t)r * nblk + b) * C);
}

// One 64-lane wavefront per replica: lane l adds partials l, l + 64, ... in
// order, then the lanes are combined by a fixed shuffle tree.
template <int CA, bool ENERGY_ONLY, class Acc>
__global__ __launch_bounds__(64) void reduce_finish_kernel(Acc acc, int R, int nblk,
                                                           const int32_t *__restrict__ tind,
                                                           const double *__restrict__ cached,
                                                           double *__restrict__ out,
                                                           const double *__restrict__ partials) {
    constexpr int C = ENERGY_ONLY ? 1 : CA;
    const int r = blockIdx.x, lane = threadIdx.x;
    double t[CA];
#pragma unroll
    for (int c = 0; c < CA; c++) t[c] = 0.0;
    const double *p = partials + (size_t)r * nblk * C;
    for (int bb = lane; bb < nblk; bb += 64) {
#pragma unroll
        for (int c = 0; c < C; c++) t[c] += p[(size_t)bb * C + c];
    }
#pragma unroll
    for (int c = 0; c < C; c++) t[c] = wave_sum(t[c]);
    if (lane != 0) return;
    if (ENERGY_ONLY) {
        out[r] = t[0];
    } else {
        const int dest = tind ? tind[r] : r;
        acc.finalize(t, cached ? cached[r] : t[0], out, R, dest);
    }
}

// Launch helper: full measurement (out is (C_out, R)) or energy (out is (R,)).
template <int CA, bool ENERGY_ONLY, class Acc>
int launch_reduce(void *stream, const Acc &acc, int R, int L, const int32_t *tind,
                  const double *cached, double *out) {
    MC_REQUIRE(R >= 1, "need at least one replica, got %d", R);
    MC_REQUIRE(L >= 1, "invalid lattice size %d", L);
    constexpr int C = ENERGY_ONLY ? 1 : CA;
    const int nblk = reduce_blocks(L);
    void *base = nullptr;
    const size_t need = (size_t)R * nblk * C;
    if (g_partials_override.ptr) {
        MC_REQUIRE(need <= g_partials_override.capacity,
                   "the reduction needs %zu partials; the caller's buffer holds %zu", need,
                   g_partials_override.capacity);
        base = g_partials_override.ptr;
    } else {
        int rc = reduction_scratch(stream, need * sizeof(double), &base);
        if (rc) return rc;
    }
    double *partials = reinterpret_cast<double *>(base);
    hipLaunchKernelGGL((reduce_kernel<CA, ENERGY_ONLY, Acc>), dim3(nblk, R), dim3(kReduceThreads), 0,
                       as_stream(stream), acc, L, nblk, cached != nullptr, partials);
    MC_CHECK_LAUNCH();
    hipLaunchKernelGGL((reduce_finish_kernel<CA, ENERGY_ONLY, Acc>), dim3(R), dim3(64), 0,
                       as_stream(stream), acc, R, nblk, tind, cached, out, partials);
    MC_CHECK_LAUNCH();
    return MC_OK;
}

// Writes t[0..C-1] unchanged, with channel 0 replaced by `energy`.
template <int C>
__device__ __forceinline__ void write_plain(const double (&t)[C], double energy, double *out, int R,
                                            int dest) {
    out[dest] = energy;
#pragma unroll
    for (int c = 1; c < C; c++) out[(size_t)c * R + dest] = t[c];
}

}  // namespace mc
