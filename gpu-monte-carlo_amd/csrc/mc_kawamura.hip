// Kawamura chiral film model (PRL 82, 964, Eq. 3) on the square lattice:
// state per site (theta, tau, dx, dy, dz); three Metropolis sub-moves per
// site visit and the seven measurement channels.
#include "mc_common.h"
#include "mc_reduce.h"

namespace mc {

// Local energy of site (i, j) from its four bonds (neighbours up, down,
// left, right) given per bond the neighbour's tau, the spin dot product and
// cos(theta - theta_n): -sum ((J (1 + tau tau_n)) dot) cos, accumulated in
// bond order as models/kawamura/update.py:7-23 does.
template <class R>
__device__ __forceinline__ R kaw_bonds(const R (&ntau)[4], const R (&dot)[4], const R (&cs)[4], R tau, R J) {
    R energy = R(0);
#pragma unroll
    for (int k = 0; k < 4; k++) energy -= ((J * (R(1) + tau * ntau[k])) * dot[k]) * cs[k];
    return energy;
}

// One site visit (models/kawamura/update.py:26-65): theta move, tau flip,
// director move, each against the energy left by the previous sub-move.  The
// neighbours (nb[q] points at neighbour q's five reals: up, down, left,
// right) are read once; cos(theta - theta_n) and the dot products are
// computed once per value of theta and of the director and reused by the
// later sub-moves (the same operations on the same operands as evaluating
// every bond afresh, so the results are the same).  `s` holds the site's
// five reals and is updated in place.  Both the per-colour kernel (global
// memory) and the sweep kernel (LDS rows) call this, so they agree bit for bit.
template <class R, class P>
__device__ __forceinline__ void kawamura_visit(R (&s)[5], const P (&nb)[4], Rng &g, R beta, R J, R delta) {
    const R one = R(1), two = R(2);
    const R pi = RealConst<R>::pi(), two_pi = RealConst<R>::two_pi();
    R theta = s[0], tau = s[1], dx = s[2], dy = s[3], dz = s[4];
    R nth[4], ntau[4], nx[4], ny[4], nz[4], dot[4], cs[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        nth[q] = nb[q][0];
        ntau[q] = nb[q][1];
        nx[q] = nb[q][2];
        ny[q] = nb[q][3];
        nz[q] = nb[q][4];
        dot[q] = (dx * nx[q] + dy * ny[q]) + dz * nz[q];
        cs[q] = rcos(theta - nth[q]);
    }
    R energy = kaw_bonds(ntau, dot, cs, tau, J);

    R pt = theta + delta * (two * g.template uniform<R>() - one);
    if (pt >= pi)
        pt -= two_pi;
    else if (pt < -pi)
        pt += two_pi;
    R pcs[4];
#pragma unroll
    for (int q = 0; q < 4; q++) pcs[q] = rcos(pt - nth[q]);
    R pe = kaw_bonds(ntau, dot, pcs, tau, J);
    R dE = pe - energy;
    R u = g.template uniform<R>();
    if (metropolis_accept(dE, u, beta)) {
        theta = pt;
        energy = pe;
#pragma unroll
        for (int q = 0; q < 4; q++) cs[q] = pcs[q];
    }

    const R ptau = -tau;
    pe = kaw_bonds(ntau, dot, cs, ptau, J);
    dE = pe - energy;
    u = g.template uniform<R>();
    if (metropolis_accept(dE, u, beta)) {
        tau = ptau;
        energy = pe;
    }

    const R z = two * g.template uniform<R>() - one;
    const R phi = two_pi * g.template uniform<R>();
    const R radius = rsqrt_(one - z * z);
    const R x = radius * rcos(phi), y = radius * rsin(phi);
    R pdot[4];
#pragma unroll
    for (int q = 0; q < 4; q++) pdot[q] = (x * nx[q] + y * ny[q]) + z * nz[q];
    pe = kaw_bonds(ntau, pdot, cs, tau, J);
    dE = pe - energy;
    u = g.template uniform<R>();
    if (metropolis_accept(dE, u, beta)) {
        dx = x;
        dy = y;
        dz = z;
    }
    s[0] = theta;
    s[1] = tau;
    s[2] = dx;
    s[3] = dy;
    s[4] = dz;
}

template <class R>
__global__ __launch_bounds__(256) void kawamura_color_kernel(R *__restrict__ st, uint64_t *__restrict__ rng,
                                                             const R *__restrict__ betas, int L,
                                                             int color, R J, R delta) {
    const int h = L >> 1;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= h * h) return;
    const int r = blockIdx.y;
    const int row = t / h, col = t - (t / h) * h;
    const int i = 2 * row + (color >> 1), j = 2 * col + (color & 1);
    R *b = st + (size_t)r * L * L * 5;
    const size_t k = (size_t)r * L * L + (size_t)i * L + j;
    R *site = b + 5 * (i * L + j);
    R s[5] = {site[0], site[1], site[2], site[3], site[4]};
    const R *nb[4] = {b + 5 * (wrap_up(i, L) * L + j), b + 5 * (wrap_dn(i, L) * L + j),
                      b + 5 * (i * L + wrap_up(j, L)), b + 5 * (i * L + wrap_dn(j, L))};
    Rng g = load_rng(rng, k);
    kawamura_visit(s, nb, g, betas[r], J, delta);
    store_rng(rng, k, g);
#pragma unroll
    for (int c = 0; c < 5; c++) site[c] = s[c];
}

// Half a sweep with rows staged in LDS: half 0 runs colours 0 then 1 (the
// sites of even rows), half 1 colours 2 then 3 (odd rows).  Colours 0 and 1
// read only odd rows besides their own row, and colours 2 and 3 only even
// rows, so a launch reads rows that no workgroup of the same launch writes:
// workgroups need no ordering among themselves, and two launches make a
// sweep.  A workgroup owns the B/2 rows i0 + half + 2k (k < B/2) of one
// replica; LDS row lr holds lattice row (i0 + half - 1 + lr) mod L,
// lr = 0..B, so its own rows are the odd lr and the rows they read the even
// lr.  Rows are loaded whole and coalesced; the own rows go back whole (both
// colours of a row change in the same launch), and thread (p, q) keeps the
// streams of its two sites (row p, columns 2q and 2q + 1) in registers.
template <class R>
__global__ __launch_bounds__(512) void kawamura_rows_kernel(R *__restrict__ st, uint64_t *__restrict__ rng, const R *__restrict__ betas,
                                     int L, int B, int half, R J, R delta) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    R *rows = reinterpret_cast<R *>(smem);
    const int r = blockIdx.y, i0 = blockIdx.x * B, h = L >> 1;
    const int tid = threadIdx.x, nt = blockDim.x;
    R *b = st + (size_t)r * L * L * 5;
    const int row_reals = L * 5;
    const int p = tid / h, q = tid - (tid / h) * h;
    const int gi = i0 + half + 2 * p, lr = 1 + 2 * p;
    const size_t k0 = (size_t)r * L * L + (size_t)gi * L + 2 * q;
    // the streams first: in flight while the rows are staged
    Rng g0 = load_rng(rng, k0), g1 = load_rng(rng, k0 + 1);
    const R beta = betas[r];
    // rows 0..B: global rows i0 + half - 1 + lr (mod L)
    stage_rows(rows, b, L, row_reals, B + 1, i0 + half - 1, tid, nt);
    __syncthreads();
#pragma unroll
    for (int cc = 0; cc < 2; cc++) {
        const int j = 2 * q + cc;
        R *site = rows + lr * row_reals + 5 * j;
        R s[5] = {site[0], site[1], site[2], site[3], site[4]};
        const R *nb[4] = {rows + (lr - 1) * row_reals + 5 * j, rows + (lr + 1) * row_reals + 5 * j,
                          rows + lr * row_reals + 5 * wrap_up(j, L), rows + lr * row_reals + 5 * wrap_dn(j, L)};
        kawamura_visit(s, nb, cc == 0 ? g0 : g1, beta, J, delta);
#pragma unroll
        for (int x = 0; x < 5; x++) site[x] = s[x];
        __syncthreads();
    }
    store_rng(rng, k0, g0);
    store_rng(rng, k0 + 1, g1);
    write_rows(b, rows, row_reals, B / 2, i0 + half - 1, tid, nt);
}

// Rows per workgroup of the row kernels: the even divisor B of L whose
// workgroup of (B/2) * (L/2) threads comes closest to 512 without exceeding
// it, with B + 1 rows of LDS within 64 KiB; 0 when none fits (the per-colour
// kernels run instead).
static int kawamura_rows_per_block(int L, size_t real_bytes) {
    int best = 0, best_gap = 1 << 30;
    for (int B = 2; B <= L; B += 2) {
        if (L % B) continue;
        const int threads = (B / 2) * (L / 2);
        const size_t lds = (size_t)(B + 1) * L * 5 * real_bytes;
        if (threads > 512 || lds > 64 * 1024) continue;
        const int gap = threads > 512 ? threads - 512 : 512 - threads;
        if (gap < best_gap) {
            best = B;
            best_gap = gap;
        }
    }
    return best;
}

// Accumulated channels: e, chirality, px, py, xx, yy, zz, xy, xz, yz, cx, qx.
template <class R>
struct KawamuraAcc {
    const R *st;
    double J;
    int L;
    template <bool ENERGY_ONLY, int CA>
    __device__ __forceinline__ void site(int r, int i, int j, double (&v)[CA], bool cached) const {
        const R *b = st + (size_t)r * L * L * 5;
        const R *s = b + 5 * (i * L + j);
        const double theta = s[0], tau = s[1], dx = s[2], dy = s[3], dz = s[4];
        if (!ENERGY_ONLY) {
            v[1] += tau;
            v[2] += ::cos(2 * theta);
            v[3] += ::sin(2 * theta);
            v[4] += dx * dx - 1.0 / 3;
            v[5] += dy * dy - 1.0 / 3;
            v[6] += dz * dz - 1.0 / 3;
            v[7] += dx * dy;
            v[8] += dx * dz;
            v[9] += dy * dz;
        }
        // With a cached energy only the x bond (twist channels) is needed.
        const int dirs = cached ? 1 : 2;
        for (int d = 0; d < dirs; d++) {
            const int ni = d == 0 ? wrap_dn(i, L) : i;
            const int nj = d == 0 ? j : wrap_dn(j, L);
            const R *n = b + 5 * (ni * L + nj);
            const double dot = (dx * (double)n[2] + dy * (double)n[3]) + dz * (double)n[4];
            const double weight = (1 + tau * (double)n[1]) * dot;
            const double angle = theta - (double)n[0];
            const double cosine = weight * ::cos(angle);
            if (!cached) v[0] -= J * cosine;
            if (!ENERGY_ONLY && d == 0) {
                v[10] += cosine;
                v[11] += weight * ::sin(angle);
            }
        }
    }
    template <int CA>
    __device__ __forceinline__ void finalize(const double (&t)[CA], double energy, double *out, int R_,
                                             int dest) const {
        out[dest] = energy;
        out[(size_t)1 * R_ + dest] = t[1];
        out[(size_t)2 * R_ + dest] = t[2];
        out[(size_t)3 * R_ + dest] = t[3];
        out[(size_t)4 * R_ + dest] =
            (t[4] * t[4] + t[5] * t[5] + t[6] * t[6] + 2 * (t[7] * t[7] + t[8] * t[8] + t[9] * t[9])) / 3;
        out[(size_t)5 * R_ + dest] = t[10];
        out[(size_t)6 * R_ + dest] = t[11];
    }
};

template <class R>
static int kawamura_update(void *stream, void *state, uint64_t *rng, const void *betas, int R_, int L,
                           int color, double J, double delta) {
    const int h = L / 2;
    dim3 grid((unsigned)((h * h + 255) / 256), (unsigned)R_);
    hipLaunchKernelGGL(kawamura_color_kernel<R>, grid, dim3(256), 0, as_stream(stream), (R *)state, rng,
                       (const R *)betas, L, color, (R)J, (R)delta);
    MC_CHECK_LAUNCH();
    return MC_OK;
}

template <class R>
static int kawamura_sweep(void *stream, void *state, uint64_t *rng, const void *betas, int R_, int L, double J,
                          double delta) {
    const int B = kawamura_rows_per_block(L, sizeof(R));
    if (B == 0) {
        for (int color = 0; color < 4; color++) {
            const int rc = kawamura_update<R>(stream, state, rng, betas, R_, L, color, J, delta);
            if (rc) return rc;
        }
        return MC_OK;
    }
    const int threads = (B / 2) * (L / 2);
    const size_t lds = (size_t)(B + 1) * L * 5 * sizeof(R);
    for (int half = 0; half < 2; half++) {
        hipLaunchKernelGGL(kawamura_rows_kernel<R>, dim3((unsigned)(L / B), (unsigned)R_), dim3(threads), lds,
                           as_stream(stream), (R *)state, rng, (const R *)betas, L, B, half, (R)J, (R)delta);
        MC_CHECK_LAUNCH();
    }
    return MC_OK;
}

}  // namespace mc

using namespace mc;

extern "C" {

int mc_kawamura_update_color(void *stream, int dtype, void *state, uint64_t *rng, const void *betas,
                             int R, int L, int color, double J, double delta) {
    MC_REQUIRE(L >= 2 && L % 2 == 0, "L must be even, got %d", L);
    MC_REQUIRE(R >= 1, "need at least one replica");
    MC_REQUIRE(color >= 0 && color < 4, "color must be in 0..3, got %d", color);
    if (dtype == MC_F32) return kawamura_update<float>(stream, state, rng, betas, R, L, color, J, delta);
    if (dtype == MC_F64) return kawamura_update<double>(stream, state, rng, betas, R, L, color, J, delta);
    set_error(MC_ERR_TYPE, "unsupported Kawamura dtype code %d", dtype);
    return MC_ERR_TYPE;
}

int mc_kawamura_sweep(void *stream, int dtype, void *state, uint64_t *rng, const void *betas, int R, int L,
                      double J, double delta) {
    MC_REQUIRE(L >= 2 && L % 2 == 0, "L must be even, got %d", L);
    MC_REQUIRE(R >= 1, "need at least one replica");
    if (dtype == MC_F32) return kawamura_sweep<float>(stream, state, rng, betas, R, L, J, delta);
    if (dtype == MC_F64) return kawamura_sweep<double>(stream, state, rng, betas, R, L, J, delta);
    set_error(MC_ERR_TYPE, "unsupported Kawamura dtype code %d", dtype);
    return MC_ERR_TYPE;
}

int mc_kawamura_measure(void *stream, int dtype, const void *state, int R, int L, double J,
                        const int32_t *temperature_indices, const double *cached_energies,
                        double *out) {
    if (dtype == MC_F32)
        return launch_reduce<12, false>(stream, KawamuraAcc<float>{(const float *)state, J, L}, R, L,
                                        temperature_indices, cached_energies, out);
    if (dtype == MC_F64)
        return launch_reduce<12, false>(stream, KawamuraAcc<double>{(const double *)state, J, L}, R, L,
                                        temperature_indices, cached_energies, out);
    set_error(MC_ERR_TYPE, "unsupported Kawamura dtype code %d", dtype);
    return MC_ERR_TYPE;
}

int mc_kawamura_energy(void *stream, int dtype, const void *state, int R, int L, double J,
                       double *out) {
    if (dtype == MC_F32)
        return launch_reduce<12, true>(stream, KawamuraAcc<float>{(const float *)state, J, L}, R, L,
                                       nullptr, nullptr, out);
    if (dtype == MC_F64)
        return launch_reduce<12, true>(stream, KawamuraAcc<double>{(const double *)state, J, L}, R, L,
                                       nullptr, nullptr, out);
    set_error(MC_ERR_TYPE, "unsupported Kawamura dtype code %d", dtype);
    return MC_ERR_TYPE;
}

}  // extern "C"
