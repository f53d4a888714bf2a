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
// neighbours are read once; cos(theta - theta_n) and the dot products are
// computed once per value of theta and of the director and reused by the
// later sub-moves (the same operations on the same operands as evaluating
// every bond afresh, so the results are the same).
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
    const R one = R(1), two = R(2), zero = R(0);
    const R pi = RealConst<R>::pi(), two_pi = RealConst<R>::two_pi();
    const R beta = betas[r];
    R *site = b + 5 * (i * L + j);
    R theta = site[0], tau = site[1], dx = site[2], dy = site[3], dz = site[4];
    const int ni[4] = {wrap_up(i, L), wrap_dn(i, L), i, i};
    const int nj[4] = {j, j, wrap_up(j, L), wrap_dn(j, L)};
    R nth[4], ntau[4], nx[4], ny[4], nz[4], dot[4], cs[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const R *n = b + 5 * (ni[q] * L + nj[q]);
        nth[q] = n[0];
        ntau[q] = n[1];
        nx[q] = n[2];
        ny[q] = n[3];
        nz[q] = n[4];
        dot[q] = (dx * nx[q] + dy * ny[q]) + dz * nz[q];
        cs[q] = rcos(theta - nth[q]);
    }
    Rng g = load_rng(rng, k);
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
    store_rng(rng, k, g);
    site[0] = theta;
    site[1] = tau;
    site[2] = dx;
    site[3] = dy;
    site[4] = dz;
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
