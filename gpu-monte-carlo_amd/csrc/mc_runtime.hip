// Runtime part of the C ABI: errors, device memory, streams, reduction
// scratch and the parallel numba-compatible RNG initialisation.
#include <cstring>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "mc_common.h"

namespace mc {

static thread_local char g_error[1024] = "";

void set_error(int code, const char *fmt, ...) {
    (void)code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

int fail_hip(hipError_t err, const char *what) {
    set_error(MC_ERR_HIP, "%s failed: %s (%d)", what, hipGetErrorString(err), (int)err);
    return MC_ERR_HIP;
}

// ---------------------------------------------------------------- scratch
// Reduction partials: one region per (device, stream).  Work queued on one
// stream runs in order, so its reductions may share a region; reductions on
// different streams (the drop-in path's null stream, torch's current stream
// of the autograd kernels, a batched run's own stream) never do.  A region
// only grows; before the old one is freed, the work queued on its stream is
// waited for.  The table is guarded by a mutex because ctypes calls drop the
// GIL.  Regions are never freed at process exit (no HIP call from a static
// destructor).
struct Scratch {
    int dev;
    void *stream;
    void *ptr;
    size_t bytes;
};
static std::mutex g_scratch_lock;
static std::vector<Scratch> *g_scratch = new std::vector<Scratch>();

thread_local PartialsOverride g_partials_override = {nullptr, 0};

int reduction_scratch(void *stream, size_t bytes, void **ptr) {
    int dev = 0;
    MC_TRY_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    Scratch *s = nullptr;
    for (auto &e : *g_scratch)
        if (e.dev == dev && e.stream == stream) s = &e;
    if (!s) {
        g_scratch->push_back(Scratch{dev, stream, nullptr, 0});
        s = &g_scratch->back();
    }
    if (s->bytes < bytes) {
        size_t want = bytes < ((size_t)1 << 20) ? ((size_t)1 << 20) : bytes;
        want = (want + 4095) & ~(size_t)4095;
        if (s->ptr) {
            // reductions still queued on this stream may use the old region
            MC_TRY_HIP(hipStreamSynchronize(as_stream(stream)));
            MC_TRY_HIP(hipFree(s->ptr));
            s->ptr = nullptr;
            s->bytes = 0;
        }
        void *p = nullptr;
        MC_TRY_HIP(hipMalloc(&p, want));
        s->ptr = p;
        s->bytes = want;
    }
    *ptr = s->ptr;
    return MC_OK;
}

// ------------------------------------------------------------- RNG tables
// xoroshiro128+ jump as a 128x128 matrix over GF(2): column b is the jump of
// the unit state with bit b set (b < 64 in s0, else in s1).  Table t holds
// jump^(2^t); stream k = product over the set bits of k applied to state 0.
static const int kJumpTables = 48;

struct Bits128 {
    uint64_t lo, hi;
};

static uint64_t host_rotl(uint64_t x, unsigned k) { return (x << k) | (x >> (64u - k)); }

static void host_next(uint64_t &s0, uint64_t &s1) {
    const uint64_t a = s0;
    uint64_t b = s1;
    b ^= a;
    s0 = host_rotl(a, 55) ^ b ^ (b << 14);
    s1 = host_rotl(b, 36);
}

static void host_jump(uint64_t &s0, uint64_t &s1) {
    static const uint64_t jump[2] = {0xbeac0467eba5facbULL, 0xd86b048b86aa9922ULL};
    uint64_t a = 0, b = 0;
    for (int i = 0; i < 2; i++)
        for (int bit = 0; bit < 64; bit++) {
            if (jump[i] & (1ULL << bit)) {
                a ^= s0;
                b ^= s1;
            }
            host_next(s0, s1);
        }
    s0 = a;
    s1 = b;
}

static Bits128 apply(const Bits128 *cols, Bits128 v) {
    Bits128 r{0, 0};
    for (int b = 0; b < 128; b++) {
        const uint64_t bit = b < 64 ? (v.lo >> b) & 1 : (v.hi >> (b - 64)) & 1;
        if (bit) {
            r.lo ^= cols[b].lo;
            r.hi ^= cols[b].hi;
        }
    }
    return r;
}

static std::vector<Bits128> build_tables() {
    std::vector<Bits128> t((size_t)kJumpTables * 128);
    for (int b = 0; b < 128; b++) {
        uint64_t s0 = b < 64 ? (1ULL << b) : 0, s1 = b < 64 ? 0 : (1ULL << (b - 64));
        host_jump(s0, s1);
        t[b] = Bits128{s0, s1};
    }
    for (int k = 1; k < kJumpTables; k++) {
        const Bits128 *prev = &t[(size_t)(k - 1) * 128];
        Bits128 *cur = &t[(size_t)k * 128];
        for (int b = 0; b < 128; b++) cur[b] = apply(prev, prev[b]);
    }
    return t;
}

static std::mutex g_table_mutex;
static std::vector<Bits128> g_host_tables;
static void *g_dev_tables[64] = {nullptr};

static int device_tables(const Bits128 **dev_ptr) {
    int dev = 0;
    MC_TRY_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_table_mutex);
    if (g_host_tables.empty()) g_host_tables = build_tables();
    if (!g_dev_tables[dev]) {
        void *p = nullptr;
        const size_t bytes = g_host_tables.size() * sizeof(Bits128);
        MC_TRY_HIP(hipMalloc(&p, bytes));
        MC_TRY_HIP(hipMemcpy(p, g_host_tables.data(), bytes, hipMemcpyHostToDevice));
        g_dev_tables[dev] = p;
    }
    *dev_ptr = reinterpret_cast<const Bits128 *>(g_dev_tables[dev]);
    return MC_OK;
}

__global__ __launch_bounds__(256) void rng_init_kernel(uint64_t *states, int64_t n, uint64_t s0,
                                                       uint64_t s1, const Bits128 *tables) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    uint64_t lo = s0, hi = s1;
    for (int t = 0; t < kJumpTables; t++) {
        if (((uint64_t)k >> t) == 0) break;
        if (((uint64_t)k >> t) & 1) {
            const Bits128 *cols = tables + (size_t)t * 128;
            uint64_t rlo = 0, rhi = 0;
#pragma unroll 4
            for (int b = 0; b < 64; b++) {
                const uint64_t m = 0 - ((lo >> b) & 1);
                rlo ^= cols[b].lo & m;
                rhi ^= cols[b].hi & m;
            }
#pragma unroll 4
            for (int b = 0; b < 64; b++) {
                const uint64_t m = 0 - ((hi >> b) & 1);
                rlo ^= cols[64 + b].lo & m;
                rhi ^= cols[64 + b].hi & m;
            }
            lo = rlo;
            hi = rhi;
        }
    }
    ulonglong2 v;
    v.x = lo;
    v.y = hi;
    reinterpret_cast<ulonglong2 *>(states)[k] = v;
}

}  // namespace mc

using namespace mc;

extern "C" {

const char *mc_last_error(void) { return g_error; }

const char *mc_version(void) { return "mcgpu 0.1 gfx950"; }

int mc_device_count(int *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return MC_OK;
}

int mc_init(int device) {
    MC_TRY_HIP(hipSetDevice(device));
    const Bits128 *t = nullptr;
    return device_tables(&t);
}

int mc_malloc(void **ptr, size_t bytes) {
    *ptr = nullptr;
    if (bytes == 0) bytes = 16;
    MC_TRY_HIP(hipMalloc(ptr, bytes));
    return MC_OK;
}

int mc_free(void *ptr) {
    if (ptr) MC_TRY_HIP(hipFree(ptr));
    return MC_OK;
}

int mc_host_alloc(void **ptr, size_t bytes) {
    *ptr = nullptr;
    if (bytes == 0) bytes = 16;
    MC_TRY_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    std::memset(*ptr, 0, bytes);
    return MC_OK;
}

int mc_host_free(void *ptr) {
    if (ptr) MC_TRY_HIP(hipHostFree(ptr));
    return MC_OK;
}

int mc_memcpy_htod(void *dst, const void *src, size_t bytes) {
    if (bytes) MC_TRY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MC_OK;
}

int mc_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
    if (bytes) MC_TRY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return MC_OK;
}

int mc_memcpy_dtod(void *dst, const void *src, size_t bytes) {
    if (bytes) MC_TRY_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return MC_OK;
}

int mc_memset(void *dst, int value, size_t bytes) {
    if (bytes) MC_TRY_HIP(hipMemset(dst, value, bytes));
    return MC_OK;
}

int mc_synchronize(void) {
    MC_TRY_HIP(hipDeviceSynchronize());
    return MC_OK;
}

int mc_stream_create(void **stream) {
    hipStream_t s = nullptr;
    MC_TRY_HIP(hipStreamCreate(&s));
    *stream = s;
    return MC_OK;
}

int mc_stream_destroy(void *stream) {
    if (stream) MC_TRY_HIP(hipStreamDestroy(as_stream(stream)));
    return MC_OK;
}

int mc_stream_synchronize(void *stream) {
    MC_TRY_HIP(hipStreamSynchronize(as_stream(stream)));
    return MC_OK;
}

int mc_event_create(void **event) {
    hipEvent_t e = nullptr;
    MC_TRY_HIP(hipEventCreate(&e));
    *event = e;
    return MC_OK;
}

int mc_event_destroy(void *event) {
    if (event) MC_TRY_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
    return MC_OK;
}

int mc_event_record(void *event, void *stream) {
    MC_TRY_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(event), as_stream(stream)));
    return MC_OK;
}

int mc_event_elapsed_ms(void *start, void *stop, float *ms) {
    MC_TRY_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
    MC_TRY_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop)));
    return MC_OK;
}

int mc_device_info(char *name, int name_bytes, int *compute_units) {
    int dev = 0;
    MC_TRY_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    MC_TRY_HIP(hipGetDeviceProperties(&prop, dev));
    if (name && name_bytes > 0) {
        if (prop.name[0])
            snprintf(name, (size_t)name_bytes, "%s (%s)", prop.name, prop.gcnArchName);
        else  // some runtimes leave the marketing name empty
            snprintf(name, (size_t)name_bytes, "%s", prop.gcnArchName);
    }
    if (compute_units) *compute_units = prop.multiProcessorCount;
    return MC_OK;
}

int mc_rng_init(void *stream, uint64_t *states, int64_t n, uint64_t seed,
                uint64_t subsequence_start) {
    MC_REQUIRE(n >= 0, "number of RNG states must be nonnegative, got %lld", (long long)n);
    if (n == 0) return MC_OK;
    MC_REQUIRE((uint64_t)n < (1ULL << kJumpTables), "too many RNG states: %lld", (long long)n);
    const Bits128 *tables = nullptr;
    int rc = device_tables(&tables);
    if (rc) return rc;
    // splitmix64(seed) -> s0 = s1 = z (numba/cuda/random.py:45-68)
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z = z ^ (z >> 31);
    uint64_t s0 = z, s1 = z;
    // advance the first state by subsequence_start jumps using the tables
    {
        Bits128 v{s0, s1};
        for (int t = 0; t < kJumpTables && (subsequence_start >> t); t++)
            if ((subsequence_start >> t) & 1) v = apply(&g_host_tables[(size_t)t * 128], v);
        MC_REQUIRE((subsequence_start >> kJumpTables) == 0, "subsequence_start too large");
        s0 = v.lo;
        s1 = v.hi;
    }
    const int64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(rng_init_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream),
                       states, n, s0, s1, tables);
    MC_CHECK_LAUNCH();
    return MC_OK;
}

}  // extern "C"
