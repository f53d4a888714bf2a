// On-device replica exchange: one matching of neighbouring temperature pairs
// swaps betas and the walker<->temperature maps (configurations never move),
// then updates the per-walker round-trip counters.
#include "mc_common.h"

namespace mc {

// rows of the trips array (5, R), indexed by walker
enum { kOrigin = 0, kVisitedOther = 1, kStartSweep = 2, kRoundTrips = 3, kRoundTripSweeps = 4 };

template <class B>
__global__ __launch_bounds__(256) void exchange_kernel(const double *__restrict__ energies, B *__restrict__ betas,
                                                       int32_t *__restrict__ w2t, int32_t *__restrict__ t2w,
                                                       uint64_t *__restrict__ pt_rng, int64_t *__restrict__ counts,
                                                       int64_t *__restrict__ trips, int R, int parity,
                                                       int64_t sweep) {
    const int tid = threadIdx.x;
    const int P = R - 1;
    // pairs of one parity are disjoint, so threads never touch the same walker
    for (int left = parity + 2 * tid; left < R - 1; left += 2 * blockDim.x) {
        const int right = left + 1;
        const int a = t2w[left], b = t2w[right];
        const double ea = energies[a], eb = energies[b];
        counts[left] += 1;
        if (!isfinite(ea) || !isfinite(eb)) {
            counts[2 * P + left] += 1;
            continue;
        }
        const double log_ratio = ((double)betas[a] - (double)betas[b]) * (ea - eb);
        Rng g = load_rng(pt_rng, left);
        const double u = g.uniform_f64();
        store_rng(pt_rng, left, g);
        if (log_ratio >= 0.0 || u == 0.0 || ::log(u) < log_ratio) {
            t2w[left] = b;
            t2w[right] = a;
            w2t[a] = right;
            w2t[b] = left;
            const B tmp = betas[a];
            betas[a] = betas[b];
            betas[b] = tmp;
            counts[P + left] += 1;
        }
    }
    // every map write precedes any endpoint check
    __syncthreads();
    for (int w = tid; w < R; w += blockDim.x) {
        const int t = w2t[w];
        if (t != 0 && t != R - 1) continue;
        const int64_t origin = trips[kOrigin * R + w];
        if (origin < 0) {
            trips[kOrigin * R + w] = t;
            trips[kStartSweep * R + w] = sweep;
        } else if (t != origin) {
            trips[kVisitedOther * R + w] = 1;
        } else if (trips[kVisitedOther * R + w]) {
            trips[kRoundTrips * R + w] += 1;
            trips[kRoundTripSweeps * R + w] += sweep - trips[kStartSweep * R + w];
            trips[kStartSweep * R + w] = sweep;
            trips[kVisitedOther * R + w] = 0;
        }
    }
}

__global__ __launch_bounds__(128) void record_walkers_kernel(const int32_t *__restrict__ w2t,
                                                             const int32_t *__restrict__ ids, int n,
                                                             int32_t *__restrict__ row) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) row[c] = w2t[ids[c]];
}

}  // namespace mc

using namespace mc;

extern "C" {

int mc_pt_exchange(void *stream, int beta_dtype, const double *energies, void *betas,
                   int32_t *walker_to_temperature, int32_t *temperature_to_walker, uint64_t *pt_rng,
                   int64_t *counts, int64_t *trips, int R, int parity, int64_t sweep) {
    MC_REQUIRE(R >= 2, "parallel tempering needs at least two replicas, got %d", R);
    MC_REQUIRE(parity == 0 || parity == 1, "parity must be 0 or 1, got %d", parity);
    if (beta_dtype == MC_F32) {
        hipLaunchKernelGGL(exchange_kernel<float>, dim3(1), dim3(256), 0, as_stream(stream), energies,
                           (float *)betas, walker_to_temperature, temperature_to_walker, pt_rng, counts,
                           trips, R, parity, sweep);
    } else if (beta_dtype == MC_F64) {
        hipLaunchKernelGGL(exchange_kernel<double>, dim3(1), dim3(256), 0, as_stream(stream), energies,
                           (double *)betas, walker_to_temperature, temperature_to_walker, pt_rng, counts,
                           trips, R, parity, sweep);
    } else {
        set_error(MC_ERR_TYPE, "unsupported beta dtype code %d", beta_dtype);
        return MC_ERR_TYPE;
    }
    MC_CHECK_LAUNCH();
    return MC_OK;
}

int mc_pt_record_walkers(void *stream, const int32_t *walker_to_temperature, const int32_t *walker_ids,
                         int n, int32_t *history_row) {
    MC_REQUIRE(n >= 1, "need at least one tracked walker");
    hipLaunchKernelGGL(record_walkers_kernel, dim3((unsigned)((n + 127) / 128)), dim3(128), 0,
                       as_stream(stream), walker_to_temperature, walker_ids, n, history_row);
    MC_CHECK_LAUNCH();
    return MC_OK;
}

}  // extern "C"
