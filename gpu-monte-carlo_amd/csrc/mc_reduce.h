// Deterministic per-replica reductions for the measurement and energy kernels.
//
// Grid (nblk, R) x 256 threads.  Each block strides over a replica's sites,
// reduces its float64 channels with wavefront shuffles and LDS, writes one
// partial per channel, and the last block of each replica (agent-scope
// acq_rel counter) sums the partials in block order and writes the outputs.
// The order of every addition is fixed by (L, nblk), so the same state
// always gives bit-identical results, and an energy-only launch reproduces
// channel 0 of the full measurement exactly.
#pragma once

#include "mc_common.h"

namespace mc {

constexpr int kReduceThreads = 256;

// Blocks per replica: at least 8 sites per thread, at most 64 blocks.
inline int reduce_blocks(int L) {
    const long sites = (long)L * L;
    long b = (sites + kReduceThreads * 8 - 1) / (kReduceThreads * 8);
    if (b < 1) b = 1;
    if (b > 64) b = 64;
    return (int)b;
}

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    return x;
}

// Acc must provide:
//   template <bool ENERGY_ONLY> __device__ void site(int r, int i, int j,
//        double (&v)[CA], bool cached) const;
//   __device__ void finalize(const double (&t)[CA], double energy,
//        double *out, int R, int dest) const;   (full measurement only)
template <int CA, bool ENERGY_ONLY, class Acc>
__global__ __launch_bounds__(kReduceThreads) void reduce_kernel(
    Acc acc, int R, int L, int nblk, const int32_t *__restrict__ tind,
    const double *__restrict__ cached, double *__restrict__ out,
    double *__restrict__ partials, unsigned *__restrict__ counters) {
    constexpr int C = ENERGY_ONLY ? 1 : CA;
    const int r = blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
    const bool use_cache = cached != nullptr;
    double v[CA];
#pragma unroll
    for (int c = 0; c < CA; c++) v[c] = 0.0;
    // sites k = b * 256 + tid + m * stride in order, with (i, j) = divmod(k, L)
    // advanced by the constant step divmod(stride, L) instead of a division
    // per site
    const int n = L * L, stride = nblk * kReduceThreads;
    const int di = stride / L, dj = stride - di * L;
    int k = b * kReduceThreads + tid;
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
    __shared__ double sh[C][kReduceThreads / 64];
    const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
    for (int c = 0; c < C; c++) {
        const double s = wave_sum(v[c]);
        if (lane == 0) sh[c][wave] = s;
    }
    __syncthreads();
    __shared__ bool is_last;
    if (tid == 0) {
        double *p = partials + ((size_t)r * nblk + b) * C;
#pragma unroll
        for (int c = 0; c < C; c++) p[c] = ((sh[c][0] + sh[c][1]) + sh[c][2]) + sh[c][3];
        const unsigned prev = __hip_atomic_fetch_add(&counters[r], 1u, __ATOMIC_ACQ_REL,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        is_last = prev == (unsigned)(nblk - 1);
    }
    __syncthreads();
    if (!is_last || tid != 0) return;
    double t[CA];
#pragma unroll
    for (int c = 0; c < CA; c++) t[c] = 0.0;
    const double *p = partials + (size_t)r * nblk * C;
    for (int bb = 0; bb < nblk; bb++) {
#pragma unroll
        for (int c = 0; c < C; c++) t[c] += p[(size_t)bb * C + c];
    }
    counters[r] = 0u;
    if (ENERGY_ONLY) {
        out[r] = t[0];
    } else {
        const int dest = tind ? tind[r] : r;
        acc.finalize(t, use_cache ? cached[r] : t[0], out, R, dest);
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
    const size_t counter_bytes = (((size_t)R * sizeof(unsigned)) + 255) & ~(size_t)255;
    // Counters live at the front of the scratch region and are kept zero.
    void *base = nullptr;
    const size_t need = (size_t)64 * 1024 * sizeof(unsigned) + (size_t)R * nblk * C * sizeof(double);
    MC_REQUIRE(counter_bytes <= (size_t)64 * 1024 * sizeof(unsigned),
               "too many replicas for one reduction launch: %d", R);
    int rc = scratch(need, &base);
    if (rc) return rc;
    unsigned *counters = reinterpret_cast<unsigned *>(base);
    double *partials = reinterpret_cast<double *>(reinterpret_cast<char *>(base) +
                                                  (size_t)64 * 1024 * sizeof(unsigned));
    hipLaunchKernelGGL((reduce_kernel<CA, ENERGY_ONLY, Acc>), dim3(nblk, R), dim3(kReduceThreads), 0,
                       as_stream(stream), acc, R, L, nblk, tind, cached, out, partials, counters);
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
