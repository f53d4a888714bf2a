// Differentiable lattice Hamiltonians of the XY and Kawamura models.
//
// With a uniform boundary twist Delta_r on the x bonds (x = axis 1, the bond
// (i, j) - (i+1, j) of the reference's twist observables, models/xy/measure.py
// and models/kawamura/measure.py) and the reference's bond weights
// (w = 1 for XY, w = (1 + tau tau')(d . d') for Kawamura):
//   Cx = sum_x-bonds w cos(a - Delta),  Cy = sum_y-bonds w cos(a),
//   Sx = sum_x-bonds w sin(a - Delta),  a = theta_site - theta_neighbour,
//   H  = -J (Cx + Cy).
// Every derivative of H with respect to J and Delta is a combination of these
// sums (dH/dDelta = -J Sx, d2H/dDelta2 = J Cx, dH/dJ = -(Cx + Cy)), and every
// derivative with respect to the state is the state gradient of
// alpha Cx + beta Cy + gamma Sx for per-replica (alpha, beta, gamma).  So three
// kernels cover first and second order:
//   * terms:    per-replica (Cx, Cy, Sx), deterministic multi-block sums;
//   * gradient: per-site state gradient of alpha Cx + beta Cy + gamma Sx;
//   * hvp:      per-site Hessian-vector product of the same combination (XY).
//
// Each workgroup stages a 16 x 32 tile of one replica plus a one-site halo in
// LDS with coalesced loads, evaluating (cos theta, sin theta) once per staged
// site; every bond's cos a and sin a then follow from the angle-difference
// identities (and the twist's rotation by -Delta, its sincos taken once per
// workgroup), so a site costs about 1.2 sincos instead of one per bond end.
// Per-site terms are evaluated in the state's precision, per-replica sums in
// float64.
#include "mc_reduce.h"

namespace mc {

template <class R>
__device__ __forceinline__ void sincos_r(R x, R &s, R &c);
// float32: the argument in turns less its nearest whole number of turns,
// t = x / (2 pi) - n with 1 / (2 pi) in two parts (one fused multiply-add
// each, so t keeps the bits that x * c_hi - n cancels), then the hardware
// sine and cosine (v_sin_f32, v_cos_f32, arguments in turns).  These
// kernels' float32 results are checked against the float64 ones, not bit for
// bit (the float64 kernels call the library sincos); the polynomial sine and
// cosine of mc_fastmath.h took about 30 instructions per site here, this pair
// 6.
template <>
__device__ __forceinline__ void sincos_r<float>(float x, float &s, float &c) {
    constexpr float kTurnHi = 0x1.45f306p-3f;  // float32(1 / (2 pi))
    constexpr float kTurnLo = 0x1.b9391p-28f;  // 1 / (2 pi) - kTurnHi
    const float n = rintf(x * kTurnHi);
    float t = fmaf(x, kTurnHi, -n);
    t = fmaf(x, kTurnLo, t);                    // turns, in [-1/2, 1/2] up to rounding
    s = __builtin_amdgcn_sinf(t);
    c = __builtin_amdgcn_cosf(t);
}
// float64: whole quarter turns removed by a three-part Cody-Waite reduction,
// r = x - n pi/2 with pi/2 = kPio2[0] + kPio2[1] + kPio2[2] in fused
// multiply-adds (each step rounds once), then the minimax polynomials of
// fdlibm's __kernel_sin and __kernel_cos (Sun Microsystems, 1993) on
// [-pi/4, pi/4], each within an ulp there, and the quadrant's swap and
// signs: about 30 float64 operations, where the library sincos took about
// 150 instructions per call and, inlined for large arguments only, cost a
// kernel an occupancy step.  The first step leaves about n kPio2[1] and rounds
// it, so the result is within about 2e-16 for |x| < 2^30 and about
// |x| 5e-27 beyond, far below the rounding of x itself (|x| 1.1e-16).  NaN,
// infinite x and |x| >= 2^50 (where n pi/2 no longer cancels x to within
// pi/4) give NaN.
template <>
__device__ __forceinline__ void sincos_r<double>(double x, double &s, double &c) {
    x = fabs(x) < 0x1p50 ? x : __builtin_nan("");
    constexpr double kTwoOverPi = 0x1.45f306dc9c883p-1;
    constexpr double kPio2[3] = {0x1.921fb544p+0, 0x1.0b4611a6p-34, 0x1.3198a2ep-69};
    constexpr double kS[6] = {-0x1.5555555555549p-3, 0x1.111111110f8a6p-7, -0x1.a01a019c161d5p-13,
                              0x1.71de357b1fe7dp-19, -0x1.ae5e68a2b9cebp-26, 0x1.5d93a5acfd57cp-33};
    constexpr double kC[6] = {0x1.555555555554cp-5, -0x1.6c16c16c15177p-10, 0x1.a01a019cb159p-16,
                              -0x1.27e4f809c52adp-22, 0x1.1ee9ebdb4b1c4p-29, -0x1.8fae9be8838d4p-37};
    const double n = rint(x * kTwoOverPi);
    double r = fma(-n, kPio2[0], x);
    r = fma(-n, kPio2[1], r);
    r = fma(-n, kPio2[2], r);
    const double z = r * r;
    const double ps = fma(r * z, fma(z, fma(z, fma(z, fma(z, fma(z, kS[5], kS[4]), kS[3]), kS[2]), kS[1]), kS[0]), r);
    const double pc =
        fma(z * z, fma(z, fma(z, fma(z, fma(z, fma(z, kC[5], kC[4]), kC[3]), kC[2]), kC[1]), kC[0]), fma(-0.5, z, 1.0));
    const int q = (int)(long long)n & 3;
    const double sv = (q & 1) ? pc : ps, cv = (q & 1) ? ps : pc;
    s = (q & 2) ? -sv : sv;
    c = ((q + 1) & 2) ? -cv : cv;
}

// Per-replica values of an optional array (NULL: 0).  The load is made from a
// selected address (the array's entry or a zero in device memory) rather than
// behind a branch: a branch with its own wait before the kernels' first vector
// loads delayed them by a memory round trip per optional array.
__device__ const double kNoValue[1] = {0.0};
__device__ __forceinline__ double replica_value(const double *p, int r) { return *(p ? p + r : kNoValue); }
__device__ __forceinline__ double twist_of(const double *twist, int r) { return replica_value(twist, r); }

constexpr int kTileRows = 16, kTileCols = 32, kTileThreads = 256;
constexpr int kHaloRows = kTileRows + 2, kHaloCols = kTileCols + 2;

// Per-replica coefficients; NULL means 0.
struct Coefs {
    const double *alpha, *beta, *gamma;
    __device__ __forceinline__ void at(int r, double &a, double &b, double &g) const {
        a = replica_value(alpha, r);
        b = replica_value(beta, r);
        g = replica_value(gamma, r);
    }
};

// LDS image of one tile and its halo: local (li, lj) = global
// ((i0 + li - 1) mod L, (j0 + lj - 1) mod L); per site (cos theta, sin theta)
// and, for Kawamura, (tau, dx, dy, dz).
template <class R, int COMP>
struct Tile {
    R c[kHaloRows][kHaloCols], s[kHaloRows][kHaloCols];
    R x[COMP == 5 ? 4 : 1][kHaloRows][kHaloCols];
};

// A thread's share of a tile's halo image in registers: the loads of all its
// sites (three at most) are issued together, and the sincos and LDS writes
// come afterwards, so the loads' latency is paid once per tile.
template <class R, int COMP>
struct TileRegs {
    static constexpr int N = kHaloRows * kHaloCols, PER = (N + kTileThreads - 1) / kTileThreads;
    R v[PER][COMP];

    // Every address first, then every load: loads issued one per branch, with
    // each address computed into the registers of the previous one, were each
    // held back until the previous load returned.  A lane past the halo (the
    // last round of 256 threads covers only part of it) reads the last halo
    // site again and store() drops it.
    __device__ __forceinline__ void load(const R *__restrict__ b, int L, int i0, int j0) {
        const R *p[PER];
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int q0 = threadIdx.x + k * kTileThreads;
            const int q = q0 < N ? q0 : N - 1;
            const int li = q / kHaloCols, lj = q - (q / kHaloCols) * kHaloCols;
            int gi = i0 + li - 1, gj = j0 + lj - 1;
            if (L >= kHaloRows) {  // gi, gj in [-1, 2L): one wrap
                gi = gi < 0 ? gi + L : (gi >= L ? gi - L : gi);
                gj = gj < 0 ? gj + L : (gj >= L ? gj - L : gj);
            } else {
                gi = (gi % L + L) % L;
                gj = (gj % L + L) % L;
            }
            p[k] = b + ((size_t)gi * L + gj) * COMP;
        }
#pragma unroll
        for (int k = 0; k < PER; k++)
#pragma unroll
            for (int m = 0; m < COMP; m++) v[k][m] = p[k][m];
    }

    __device__ __forceinline__ void store(Tile<R, COMP> &t) const {
#pragma unroll
        for (int k = 0; k < PER; k++) {
            const int q = threadIdx.x + k * kTileThreads;
            if (q < N) {
                const int li = q / kHaloCols, lj = q - (q / kHaloCols) * kHaloCols;
                R sv, cv;
                sincos_r<R>(v[k][0], sv, cv);
                t.c[li][lj] = cv;
                t.s[li][lj] = sv;
                if (COMP == 5) {
#pragma unroll
                    for (int m = 0; m < 4; m++) t.x[m][li][lj] = v[k][1 + m];
                }
            }
        }
    }
};

template <class R, int COMP>
__device__ __forceinline__ void stage_tile(Tile<R, COMP> &t, const R *__restrict__ b, int L, int i0, int j0) {
    TileRegs<R, COMP> g;
    g.load(b, L, i0, j0);
    g.store(t);
}

// Bond from site p to site q (a = theta_p - theta_q), rotated by -dl:
// (cos(a - dl), sin(a - dl)) from the staged cos/sin and (cd, sd) = (cos dl, sin dl).
template <class R>
__device__ __forceinline__ void bond_cs(R cp, R sp, R cq, R sq, R cd, R sd, R &c, R &s) {
    const R ca = cp * cq + sp * sq, sa = sp * cq - cp * sq;
    c = ca * cd + sa * sd;
    s = sa * cd - ca * sd;
}

// Unrotated bond (no twist): cos a only
template <class R>
__device__ __forceinline__ R bond_c(R cp, R sp, R cq, R sq) {
    return cp * cq + sp * sq;
}

template <class R>
__device__ __forceinline__ void twist_cs(const double *twist, int r, R &cd, R &sd) {
    sincos_r<R>((R)twist_of(twist, r), sd, cd);
}

// Tiles of a replica: row-major over ceil(L / 16) x ceil(L / 32)
__device__ __forceinline__ void tile_origin(int L, int &i0, int &j0) {
    const int tc = (L + kTileCols - 1) / kTileCols;
    i0 = (blockIdx.x / tc) * kTileRows;
    j0 = (blockIdx.x - (blockIdx.x / tc) * tc) * kTileCols;
}

// ------------------------------------------------------------------ terms

// Per-replica (Cx, Cy, Sx): each workgroup sums the down (x) and right (y)
// bonds of kTermsTiles consecutive tiles (row-major tile order) into a
// float64 partial per channel; reduce_finish_kernel adds the partials in a
// fixed order (the same result for the same state every time).
constexpr int kTermsTiles = 4;

template <class R, int COMP>
__global__ __launch_bounds__(kTileThreads) void terms_tile_kernel(const R *__restrict__ st, int L,
                                                                  const double *twist, int ntiles, int nblk,
                                                                  double *__restrict__ partials) {
    __shared__ Tile<R, COMP> t;
    const int r = blockIdx.y, tid = threadIdx.x;
    const R *b = st + (size_t)r * L * L * COMP;
    const int tc = (L + kTileCols - 1) / kTileCols;
    R cd, sd;
    twist_cs<R>(twist, r, cd, sd);
    double v[3] = {0.0, 0.0, 0.0};  // float64 across tiles; each tile's few sites are summed in R
    // the next tile's loads are in flight while this one's bonds are summed
    const int first = blockIdx.x * kTermsTiles, last = min(ntiles, first + kTermsTiles);
    TileRegs<R, COMP> g;
    if (first < last) g.load(b, L, (first / tc) * kTileRows, (first - (first / tc) * tc) * kTileCols);
    for (int tile = first; tile < last; tile++) {
        const int i0 = (tile / tc) * kTileRows, j0 = (tile - (tile / tc) * tc) * kTileCols;
        __syncthreads();  // the previous tile's reads are done
        g.store(t);
        if (tile + 1 < last)
            g.load(b, L, ((tile + 1) / tc) * kTileRows, ((tile + 1) - ((tile + 1) / tc) * tc) * kTileCols);
        __syncthreads();
        R tv[3] = {R(0), R(0), R(0)};
        for (int q = tid; q < kTileRows * kTileCols; q += kTileThreads) {
            const int li = 1 + q / kTileCols, lj = 1 + (q - (q / kTileCols) * kTileCols);
            if (i0 + li - 1 >= L || j0 + lj - 1 >= L) continue;
            R wx = 1, wy = 1;
            if (COMP == 5) {
                wx = ((R)1 + t.x[0][li][lj] * t.x[0][li + 1][lj]) *
                     ((t.x[1][li][lj] * t.x[1][li + 1][lj] + t.x[2][li][lj] * t.x[2][li + 1][lj]) +
                      t.x[3][li][lj] * t.x[3][li + 1][lj]);
                wy = ((R)1 + t.x[0][li][lj] * t.x[0][li][lj + 1]) *
                     ((t.x[1][li][lj] * t.x[1][li][lj + 1] + t.x[2][li][lj] * t.x[2][li][lj + 1]) +
                      t.x[3][li][lj] * t.x[3][li][lj + 1]);
            }
            R cx, sx;
            bond_cs<R>(t.c[li][lj], t.s[li][lj], t.c[li + 1][lj], t.s[li + 1][lj], cd, sd, cx, sx);
            const R cy = bond_c<R>(t.c[li][lj], t.s[li][lj], t.c[li][lj + 1], t.s[li][lj + 1]);
            tv[0] += wx * cx;
            tv[1] += wy * cy;
            tv[2] += wx * sx;
        }
#pragma unroll
        for (int c = 0; c < 3; c++) v[c] += (double)tv[c];
    }
    block_partial<3, kTileThreads>(v, partials + ((size_t)r * nblk + blockIdx.x) * 3);
}

// The per-replica channels as the outputs (3, R)
struct TermsOut {
    template <int CA>
    __device__ __forceinline__ void finalize(const double (&t)[CA], double, double *out, int R_, int dest) const {
#pragma unroll
        for (int c = 0; c < CA; c++) out[(size_t)c * R_ + dest] = t[c];
    }
};

// ------------------------------------------------------------- gradient

template <class R>
__global__ __launch_bounds__(kTileThreads) void xy_grad_kernel(const R *__restrict__ th, int L, const double *twist,
                                                               Coefs k, R *__restrict__ grad) {
    __shared__ Tile<R, 1> t;
    const int r = blockIdx.y;
    int i0, j0;
    tile_origin(L, i0, j0);
    stage_tile<R, 1>(t, th + (size_t)r * L * L, L, i0, j0);
    R cd, sd;
    twist_cs<R>(twist, r, cd, sd);
    double al, be, ga;
    k.at(r, al, be, ga);
    __syncthreads();
    for (int q = threadIdx.x; q < kTileRows * kTileCols; q += kTileThreads) {
        const int li = 1 + q / kTileCols, lj = 1 + (q - (q / kTileCols) * kTileCols);
        const int i = i0 + li - 1, j = j0 + lj - 1;
        if (i >= L || j >= L) continue;
        const R c0 = t.c[li][lj], s0 = t.s[li][lj];
        R c1, s1, c2, s2, c3, s3, c4, s4;
        bond_cs<R>(c0, s0, t.c[li + 1][lj], t.s[li + 1][lj], cd, sd, c1, s1);  // (i,j)-(i+1,j), site first
        bond_cs<R>(t.c[li - 1][lj], t.s[li - 1][lj], c0, s0, cd, sd, c2, s2);  // (i-1,j)-(i,j), site second
        bond_cs<R>(c0, s0, t.c[li][lj + 1], t.s[li][lj + 1], (R)1, (R)0, c3, s3);  // (i,j)-(i,j+1)
        bond_cs<R>(t.c[li][lj - 1], t.s[li][lj - 1], c0, s0, (R)1, (R)0, c4, s4);  // (i,j-1)-(i,j)
        // dCx/dt = -s1 + s2, dCy/dt = -s3 + s4, dSx/dt = c1 - c2
        const R g = (R)al * (s2 - s1) + (R)be * (s4 - s3) + (R)ga * (c1 - c2);
        grad[(size_t)r * L * L + (size_t)i * L + j] = g;
    }
}

template <class R>
__global__ __launch_bounds__(kTileThreads) void xy_hvp_kernel(const R *__restrict__ th, int L, const double *twist,
                                                              Coefs k, const R *__restrict__ u, R *__restrict__ out) {
    __shared__ Tile<R, 1> t;
    const int r = blockIdx.y;
    int i0, j0;
    tile_origin(L, i0, j0);
    const size_t base = (size_t)r * L * L;
    stage_tile<R, 1>(t, th + base, L, i0, j0);
    R cd, sd;
    twist_cs<R>(twist, r, cd, sd);
    double al, be, ga;
    k.at(r, al, be, ga);
    const R *v = u + base;
    __syncthreads();
    for (int q = threadIdx.x; q < kTileRows * kTileCols; q += kTileThreads) {
        const int li = 1 + q / kTileCols, lj = 1 + (q - (q / kTileCols) * kTileCols);
        const int i = i0 + li - 1, j = j0 + lj - 1;
        if (i >= L || j >= L) continue;
        const int site = i * L + j;
        const int kdn = wrap_dn(i, L) * L + j, kup = wrap_up(i, L) * L + j;
        const int krt = i * L + wrap_dn(j, L), klf = i * L + wrap_up(j, L);
        const R c0 = t.c[li][lj], s0 = t.s[li][lj];
        R c1, s1, c2, s2, c3, s3, c4, s4;
        bond_cs<R>(c0, s0, t.c[li + 1][lj], t.s[li + 1][lj], cd, sd, c1, s1);
        bond_cs<R>(t.c[li - 1][lj], t.s[li - 1][lj], c0, s0, cd, sd, c2, s2);
        bond_cs<R>(c0, s0, t.c[li][lj + 1], t.s[li][lj + 1], (R)1, (R)0, c3, s3);
        bond_cs<R>(t.c[li][lj - 1], t.s[li][lj - 1], c0, s0, (R)1, (R)0, c4, s4);
        const R vs = v[site];
        const R du_dn = vs - v[kdn], du_up = vs - v[kup];
        const R du_rt = vs - v[krt], du_lf = vs - v[klf];
        // Hess(Cx) u = -c1 du_dn - c2 du_up ; Hess(Cy) u = -c3 du_rt - c4 du_lf ;
        // Hess(Sx) u = -s1 du_dn - s2 du_up
        const R hx = -(c1 * du_dn) - c2 * du_up;
        const R hy = -(c3 * du_rt) - c4 * du_lf;
        const R hs = -(s1 * du_dn) - s2 * du_up;
        out[base + site] = (R)al * hx + (R)be * hy + (R)ga * hs;
    }
}

// Kawamura: state (theta, tau, dx, dy, dz) per site.  For a bond term
// f(a) w with w = (1 + tau_p tau_q)(d_p . d_q), a = theta_p - theta_q - Delta_b:
//   d/dtheta_p = w f'(a), d/dtheta_q = -w f'(a),
//   d/dtau_p = tau_q (d_p . d_q) f(a), d/dd_p = (1 + tau_p tau_q) f(a) d_q.
template <class R>
__global__ __launch_bounds__(kTileThreads) void kawamura_grad_kernel(const R *__restrict__ st, int L,
                                                                     const double *twist, Coefs k,
                                                                     R *__restrict__ grad) {
    __shared__ Tile<R, 5> t;
    const int r = blockIdx.y;
    int i0, j0;
    tile_origin(L, i0, j0);
    stage_tile<R, 5>(t, st + (size_t)r * L * L * 5, L, i0, j0);
    R cd, sd;
    twist_cs<R>(twist, r, cd, sd);
    double al, be, ga;
    k.at(r, al, be, ga);
    __syncthreads();
    for (int q = threadIdx.x; q < kTileRows * kTileCols; q += kTileThreads) {
        const int li = 1 + q / kTileCols, lj = 1 + (q - (q / kTileCols) * kTileCols);
        const int i = i0 + li - 1, j = j0 + lj - 1;
        if (i >= L || j >= L) continue;
        const R c0 = t.c[li][lj], s0 = t.s[li][lj];
        const R tau = t.x[0][li][lj], dx = t.x[1][li][lj], dy = t.x[2][li][lj], dz = t.x[3][li][lj];
        R g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0;
        // neighbours down, up (x bonds), right, left (y bonds); the site is
        // the bond's first end for down and right
        const int oi[4] = {li + 1, li - 1, li, li};
        const int oj[4] = {lj, lj, lj + 1, lj - 1};
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const bool xb = m < 2, first = (m & 1) == 0;
            const int a = oi[m], b = oj[m];
            const R cq = t.c[a][b], sq = t.s[a][b];
            R c, s;
            if (first)
                bond_cs<R>(c0, s0, cq, sq, xb ? cd : (R)1, xb ? sd : (R)0, c, s);
            else
                bond_cs<R>(cq, sq, c0, s0, xb ? cd : (R)1, xb ? sd : (R)0, c, s);
            const R otau = t.x[0][a][b], ox = t.x[1][a][b], oy = t.x[2][a][b], oz = t.x[3][a][b];
            const R dot = (dx * ox + dy * oy) + dz * oz;
            const R tt = (R)1 + tau * otau;
            const R w = tt * dot;
            // f = cos with coefficient (alpha | beta), f = sin with gamma (x bonds only)
            const R kc = (R)(xb ? al : be), ks = xb ? (R)ga : (R)0;
            const R f = kc * c + ks * s;   // coefficient-weighted f(a)
            const R fp = ks * c - kc * s;  // coefficient-weighted f'(a)
            g0 += first ? w * fp : -(w * fp);
            g1 += otau * dot * f;
            g2 += tt * f * ox;
            g3 += tt * f * oy;
            g4 += tt * f * oz;
        }
        R *o = grad + ((size_t)r * L * L + (size_t)i * L + j) * 5;
        o[0] = g0;
        o[1] = g1;
        o[2] = g2;
        o[3] = g3;
        o[4] = g4;
    }
}

// ------------------------------------------------ XY, whole rows per workgroup
//
// For L a multiple of 64 up to 1024 the XY kernels take whole rows: one
// workgroup of L threads per segment of kSegRows rows of one replica, thread
// j holding column j of the segment (plus the rows above and below it) in
// registers.  A site's bonds along the column are then in the thread's own
// registers and its bonds along the row come from the neighbouring lanes by
// DPP wave shifts, with each wavefront's first and last lanes exchanged through LDS
// (the wavefronts of a workgroup cover a whole row, so the row's periodic
// wrap is the last wavefront's neighbour being the first).  A site costs
// (kSegRows + 1) / kSegRows sincos (terms) or (kSegRows + 2) / kSegRows
// (gradient), against 1.2 for a 16 x 32 tile and its halo, and no LDS image
// of the tile.  The per-bond arithmetic is the tile kernels', so the gradient
// is the same bit for bit; the terms are summed per thread over its column
// segment, with the twist applied to the segment's sums of cos a and sin a
// (cos(a - D) = cos a cos D + sin a sin D is exact as an identity; the
// rounding differs from rotating each bond).
constexpr int kSegRows = 16;

// Lane i takes lane i + 1's value (shfl_dn1) or lane i - 1's (shfl_up1): a
// DPP wave shift (wave_shl:1 / wave_shr:1, one VALU instruction per 32 bits,
// checked on MI355X), the wavefront's last or first lane receiving 0 (the
// callers take those from LDS).
template <int CTRL>
__device__ __forceinline__ float wave_shift(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double wave_shift(double x) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xf, 0xf, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <class R>
__device__ __forceinline__ R shfl_dn1(R x) { return wave_shift<0x130>(x); }
template <class R>
__device__ __forceinline__ R shfl_up1(R x) { return wave_shift<0x138>(x); }

// acc + x[lane + 1] * x[lane], and acc unchanged in lane 63 (the shift
// brings 0 there): one v_fmac_f32 whose first operand is read through a DPP
// wave shift (wave_shl:1).  The compiler does not fold a separate DPP move
// into v_fmac_f32, so it is written here; s_nop 1 gives the DPP read of x
// the two wait states it needs after the VALU instruction that wrote x.
__device__ __forceinline__ float fmac_right(float acc, float x) {
    asm("s_nop 1\n\tv_fmac_f32_dpp %0, %1, %1 wave_shl:1 row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : "+v"(acc)
        : "v"(x));
    return acc;
}
__device__ __forceinline__ double fmac_right(double acc, double x) { return fma(shfl_dn1(x), x, acc); }
__device__ __forceinline__ float fma_r(float a, float b, float c) { return fmaf(a, b, c); }
__device__ __forceinline__ double fma_r(double a, double b, double c) { return fma(a, b, c); }

// Terms: each thread sums its column segment's bonds with fused
// multiply-adds; a row bond whose right end is in the next wavefront (lane
// 63's) is left out of the row loop (the wave shift brings 0), and after one
// barrier wavefront 0 adds those W * kSegRows bonds from the edge columns
// that lanes 0 and 63 left in LDS.  Every wavefront writes its own partial
// (no second barrier), so a replica has (L / kSegRows) * W partials.
template <class R, int NT>
__global__ __launch_bounds__(NT) void xy_terms_rows_kernel(const R *__restrict__ th, const double *twist,
                                                          double *__restrict__ partials) {
    constexpr int L = NT, W = NT / 64;
    __shared__ R edge[2][2][W][kSegRows];  // [lane 0, lane 63][cos, sin][wavefront][row]
    const int r = blockIdx.y, j = threadIdx.x, w = j >> 6, lane = j & 63;
    const int i0 = blockIdx.x * kSegRows;
    const R *b = th + (size_t)r * L * L + j;
    R v[kSegRows + 1], c[kSegRows + 1], sn[kSegRows + 1];
#pragma unroll
    for (int k = 0; k <= kSegRows; k++) v[k] = b[(size_t)((i0 + k) & (L - 1)) * L];
    R cd, sd;
    twist_cs<R>(twist, r, cd, sd);
#pragma unroll
    for (int k = 0; k <= kSegRows; k++) sincos_r<R>(v[k], sn[k], c[k]);
    R ca = 0, sa = 0, cy = 0;
#pragma unroll
    for (int k = 0; k < kSegRows; k++) {
        ca = fma_r(c[k], c[k + 1], fma_r(sn[k], sn[k + 1], ca));   // x bond (i, j) - (i+1, j): cos a
        sa = fma_r(sn[k], c[k + 1], fma_r(-c[k], sn[k + 1], sa));  //                          sin a
        cy = fmac_right(fmac_right(cy, c[k]), sn[k]);              // y bond (i, j) - (i, j+1)
    }
    if (lane == 0 || lane == 63) {
        const int e = lane == 0 ? 0 : 1;
#pragma unroll
        for (int k = 0; k < kSegRows; k++) {
            edge[e][0][w][k] = c[k];
            edge[e][1][w][k] = sn[k];
        }
    }
    __syncthreads();
    if (w == 0)
        for (int q = lane; q < W * kSegRows; q += 64) {
            const int ww = q / kSegRows, k = q - ww * kSegRows, wn = (ww + 1) % W;
            cy = fma_r(edge[1][0][ww][k], edge[0][0][wn][k], cy);
            cy = fma_r(edge[1][1][ww][k], edge[0][1][wn][k], cy);
        }
    // this thread's x bonds rotated by -Delta as sums
    const double vv[3] = {(double)(ca * cd + sa * sd), (double)cy, (double)(sa * cd - ca * sd)};
    double *o = partials + (((size_t)r * gridDim.x + blockIdx.x) * W + w) * 3;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double t = wave_sum(vv[q]);
        if (lane == 0) o[q] = t;
    }
}

// The twisted x bond p - q: (cos(a - D), sin(a - D)) with fused
// multiply-adds (the whole-row kernels' rounding, not the tile kernels').
template <class R>
__device__ __forceinline__ void bond_x_fma(R cp, R sp, R cq, R sq, R cd, R sd, R &c, R &s) {
    const R ca = fma_r(cp, cq, sp * sq), sa = fma_r(sp, cq, -(cp * sq));
    c = fma_r(ca, cd, sa * sd);
    s = fma_r(sa, cd, -(ca * sd));
}

// Gradient: every bond is evaluated once.  The x bond below row k is the
// site's down bond (s1, c1) and, carried to row k + 1, that site's up bond
// (s2, c2); the y bond to the right is the site's s3, and the left
// neighbour's s3, one lane over, is its s4.  A wavefront's lane 63 takes its
// right neighbour, and lane 0 its left neighbour's y bond, from the edge
// columns in LDS.
template <class R, int NT>
__global__ __launch_bounds__(NT) void xy_grad_rows_kernel(const R *__restrict__ th, const double *twist, Coefs kc,
                                                         R *__restrict__ grad) {
    constexpr int L = NT, W = NT / 64;
    __shared__ R edge[2][2][W][kSegRows];  // [lane 0, lane 63][cos, sin][wavefront][row]
    const int r = blockIdx.y, j = threadIdx.x, w = j >> 6, lane = j & 63;
    const int i0 = blockIdx.x * kSegRows;
    const size_t base = (size_t)r * L * L;
    const R *b = th + base + j;
    // rows i0 - 1 .. i0 + kSegRows (index k + 1 holds row i0 + k)
    R v[kSegRows + 2], c[kSegRows + 2], sn[kSegRows + 2];
#pragma unroll
    for (int k = 0; k < kSegRows + 2; k++) v[k] = b[(size_t)((i0 + k - 1) & (L - 1)) * L];
    R cd, sd;
    twist_cs<R>(twist, r, cd, sd);
    double al, be, ga;
    kc.at(r, al, be, ga);
#pragma unroll
    for (int k = 0; k < kSegRows + 2; k++) sincos_r<R>(v[k], sn[k], c[k]);
    if (lane == 0 || lane == 63) {
        const int e = lane == 0 ? 0 : 1;
#pragma unroll
        for (int k = 0; k < kSegRows; k++) {
            edge[e][0][w][k] = c[k + 1];
            edge[e][1][w][k] = sn[k + 1];
        }
    }
    __syncthreads();
    const int wn = (w + 1) % W, wp = (w + W - 1) % W;
    const R a = (R)al, bb = (R)be, g_ = (R)ga;
    R c2, s2;  // the up bond (i-1, j) - (i, j)
    bond_x_fma<R>(c[0], sn[0], c[1], sn[1], cd, sd, c2, s2);
#pragma unroll
    for (int k = 1; k <= kSegRows; k++) {
        const R c0 = c[k], s0 = sn[k];
        R c1, s1;  // the down bond (i, j) - (i+1, j)
        bond_x_fma<R>(c0, s0, c[k + 1], sn[k + 1], cd, sd, c1, s1);
        R cr = shfl_dn1(c0), sr = shfl_dn1(s0);
        if (lane == 63) {
            cr = edge[0][0][wn][k - 1];
            sr = edge[0][1][wn][k - 1];
        }
        const R s3 = fma_r(s0, cr, -(c0 * sr));  // (i, j) - (i, j+1)
        R s4 = shfl_up1(s3);                     // (i, j-1) - (i, j)
        if (lane == 0) s4 = fma_r(edge[1][1][wp][k - 1], c0, -(edge[1][0][wp][k - 1] * s0));
        // dCx/dt = -s1 + s2, dCy/dt = -s3 + s4, dSx/dt = c1 - c2
        const R g = fma_r(a, s2 - s1, fma_r(bb, s4 - s3, g_ * (c1 - c2)));
        grad[base + (size_t)(i0 + k - 1) * L + j] = g;
        c2 = c1;
        s2 = s1;
    }
}

// ------------------------------------------- XY, one wavefront per row band
//
// For L = 64, 128, 256 a wavefront covers whole rows: lane l holds the M =
// L / 64 adjacent columns M l .. M l + M - 1 (one vector load per row) of a
// band of BandRows<R> rows (16 in float32, 8 in float64, whose registers are
// twice as wide).  A site's right neighbour is in the same lane, or (last
// column) lane l + 1's first column by a DPP wave rotate, which also closes
// the periodic row (lane 63 takes lane 0's): no LDS, no barrier, no edge
// fix-ups, and one wavefront per workgroup.  Each bond is evaluated once: the
// x bond below row k is the site's down bond and the next row's up bond; the
// y bond to the right is the site's s3 and its right neighbour's s4.
template <class R>
struct BandRows {
    static constexpr int value = sizeof(R) == 4 ? 16 : 8;
};

template <int CTRL>
__device__ __forceinline__ float wave_rotate(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double wave_rotate(double x) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// lane i takes lane i + 1's value, lane 63 lane 0's (wave_rol:1)
template <class R>
__device__ __forceinline__ R rot_next(R x) { return wave_rotate<0x134>(x); }
// lane i takes lane i - 1's value, lane 0 lane 63's (wave_ror:1)
template <class R>
__device__ __forceinline__ R rot_prev(R x) { return wave_rotate<0x13C>(x); }

// M adjacent values in 16-byte (or smaller) vector accesses
template <class R, int M>
__device__ __forceinline__ void load_cols(const R *p, R (&x)[M]) {
    constexpr int W = 16 / (int)sizeof(R);  // values per 16-byte access
    if constexpr (M >= W) {
#pragma unroll
        for (int q = 0; q < M; q += W) {
            if constexpr (sizeof(R) == 4) {
                const float4 v = *reinterpret_cast<const float4 *>(p + q);
                x[q] = v.x; x[q + 1] = v.y; x[q + 2] = v.z; x[q + 3] = v.w;
            } else {
                const double2 v = *reinterpret_cast<const double2 *>(p + q);
                x[q] = v.x; x[q + 1] = v.y;
            }
        }
    } else if constexpr (M == 2) {
        const float2 v = *reinterpret_cast<const float2 *>(p);
        x[0] = v.x; x[1] = v.y;
    } else {
        x[0] = *p;
    }
}
template <class R, int M>
__device__ __forceinline__ void store_cols(R *p, const R (&x)[M]) {
    constexpr int W = 16 / (int)sizeof(R);
    if constexpr (M >= W) {
#pragma unroll
        for (int q = 0; q < M; q += W) {
            if constexpr (sizeof(R) == 4)
                *reinterpret_cast<float4 *>(p + q) = make_float4(x[q], x[q + 1], x[q + 2], x[q + 3]);
            else
                *reinterpret_cast<double2 *>(p + q) = make_double2(x[q], x[q + 1]);
        }
    } else if constexpr (M == 2) {
        *reinterpret_cast<float2 *>(p) = make_float2(x[0], x[1]);
    } else {
        *p = x[0];
    }
}

// Terms: per-wavefront partials (Cx, Cy, Sx), BandRows<R> rows each.
template <class R, int M>
__global__ __launch_bounds__(64) void xy_terms_band_kernel(const R *__restrict__ th, const double *twist,
                                                          double *__restrict__ partials) {
    constexpr int L = 64 * M, BR = BandRows<R>::value;
    const int r = blockIdx.y, lane = threadIdx.x;
    const int i0 = blockIdx.x * BR;
    const R *b = th + (size_t)r * L * L + lane * M;
    R v[BR + 1][M];
#pragma unroll
    for (int k = 0; k <= BR; k++) load_cols<R, M>(b + (size_t)((i0 + k) & (L - 1)) * L, v[k]);
    R cd, sd;
    twist_cs<R>(twist, r, cd, sd);
    R cc[M], sc[M];
#pragma unroll
    for (int m = 0; m < M; m++) sincos_r<R>(v[0][m], sc[m], cc[m]);
    // one set of sums per column: with one wavefront per SIMD nothing hides the
    // latency of a dependent chain, and M chains of 2 BR terms are shorter
    // than one of 2 M BR
    R ca[M] = {}, sa[M] = {}, cy[M] = {};
#pragma unroll
    for (int k = 0; k < BR; k++) {
        R cn[M], sn[M];
#pragma unroll
        for (int m = 0; m < M; m++) sincos_r<R>(v[k + 1][m], sn[m], cn[m]);
#pragma unroll
        for (int m = 0; m < M; m++) {
            ca[m] = fma_r(cc[m], cn[m], fma_r(sc[m], sn[m], ca[m]));   // x bond (i, j) - (i+1, j): cos a
            sa[m] = fma_r(sc[m], cn[m], fma_r(-cc[m], sn[m], sa[m]));  //                          sin a
            const R cr = m < M - 1 ? cc[m + 1] : rot_next(cc[0]);
            const R sr = m < M - 1 ? sc[m + 1] : rot_next(sc[0]);
            cy[m] = fma_r(cc[m], cr, fma_r(sc[m], sr, cy[m]));          // y bond (i, j) - (i, j+1)
        }
#pragma unroll
        for (int m = 0; m < M; m++) {
            cc[m] = cn[m];
            sc[m] = sn[m];
        }
    }
    R ca_t = ca[0], sa_t = sa[0], cy_t = cy[0];
#pragma unroll
    for (int m = 1; m < M; m++) {
        ca_t += ca[m];
        sa_t += sa[m];
        cy_t += cy[m];
    }
    // this lane's x bonds rotated by -Delta as sums
    const double f[3] = {(double)(ca_t * cd + sa_t * sd), (double)cy_t, (double)(sa_t * cd - ca_t * sd)};
    double *o = partials + ((size_t)r * gridDim.x + blockIdx.x) * 3;
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const double t = wave_sum(f[q]);
        if (lane == 0) o[q] = t;
    }
}

// The rows of a band: s1 (and c1) of the down x bond, carried to the next
// row as s2 (and c2); TWIST false drops the rotation by -Delta and the gamma
// term (zero twist and gamma), where they change nothing, bit for bit:
// a rotation by (1, 0) and gamma * (c1 - c2) = 0 are exact.
template <class R, int M, bool TWIST>
__device__ __forceinline__ void grad_band_rows(const R (&v)[BandRows<R>::value + 2][M], R (&cc)[M], R (&sc)[M],
                                               R (&c2)[M], R (&s2)[M], R cd, R sd, R a, R bb, R gg,
                                               R *__restrict__ out, int L) {
    constexpr int BR = BandRows<R>::value;
#pragma unroll
    for (int k = 1; k <= BR; k++) {
        R cn[M], sn[M];
#pragma unroll
        for (int m = 0; m < M; m++) sincos_r<R>(v[k + 1][m], sn[m], cn[m]);
        R s3[M];  // y bond (i, j) - (i, j+1)
#pragma unroll
        for (int m = 0; m < M - 1; m++) s3[m] = fma_r(sc[m], cc[m + 1], -(cc[m] * sc[m + 1]));
        s3[M - 1] = fma_r(sc[M - 1], rot_next(cc[0]), -(cc[M - 1] * rot_next(sc[0])));
        const R s4_first = rot_prev(s3[M - 1]);  // (i, j-1) - (i, j) of the lane's first column
        R g[M];
#pragma unroll
        for (int m = 0; m < M; m++) {
            const R s4 = m == 0 ? s4_first : s3[m - 1];
            // dCx/dt = -s1 + s2, dCy/dt = -s3 + s4, dSx/dt = c1 - c2
            if (TWIST) {
                R c1, s1;  // down bond (i, j) - (i+1, j)
                bond_x_fma<R>(cc[m], sc[m], cn[m], sn[m], cd, sd, c1, s1);
                g[m] = fma_r(a, s2[m] - s1, fma_r(bb, s4 - s3[m], gg * (c1 - c2[m])));
                c2[m] = c1;
                s2[m] = s1;
            } else {
                const R s1 = fma_r(sc[m], cn[m], -(cc[m] * sn[m]));
                g[m] = fma_r(a, s2[m] - s1, bb * (s4 - s3[m]));
                s2[m] = s1;
            }
        }
        store_cols<R, M>(out + (size_t)(k - 1) * L, g);
#pragma unroll
        for (int m = 0; m < M; m++) {
            cc[m] = cn[m];
            sc[m] = sn[m];
        }
    }
}

// Gradient of alpha Cx + beta Cy + gamma Sx, BandRows<R> rows per wavefront.
template <class R, int M>
__global__ __launch_bounds__(64) void xy_grad_band_kernel(const R *__restrict__ th, const double *twist, Coefs kc,
                                                         R *__restrict__ grad) {
    constexpr int L = 64 * M, BR = BandRows<R>::value;
    const int r = blockIdx.y, lane = threadIdx.x;
    const int i0 = blockIdx.x * BR;
    const size_t base = (size_t)r * L * L;
    const R *b = th + base + lane * M;
    // rows i0 - 1 .. i0 + BR (index k + 1 holds row i0 + k)
    R v[BR + 2][M];
#pragma unroll
    for (int k = 0; k < BR + 2; k++) load_cols<R, M>(b + (size_t)((i0 + k - 1) & (L - 1)) * L, v[k]);
    R cd, sd;
    twist_cs<R>(twist, r, cd, sd);
    double al, be, ga;
    kc.at(r, al, be, ga);
    const R a = (R)al, bb = (R)be, gg = (R)ga;
    R cc[M], sc[M], c2[M], s2[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        R cu, su;  // row i0 - 1
        sincos_r<R>(v[0][m], su, cu);
        sincos_r<R>(v[1][m], sc[m], cc[m]);
        bond_x_fma<R>(cu, su, cc[m], sc[m], cd, sd, c2[m], s2[m]);  // up bond (i-1, j) - (i, j)
    }
    R *out = grad + base + (size_t)i0 * L + lane * M;
    if (cd == (R)1 && sd == (R)0 && gg == (R)0)  // the replica's values: uniform in the wavefront
        grad_band_rows<R, M, false>(v, cc, sc, c2, s2, cd, sd, a, bb, gg, out, L);
    else
        grad_band_rows<R, M, true>(v, cc, sc, c2, s2, cd, sd, a, bb, gg, out, L);
}

// lattices of L = 64, 128, 256 take the band kernels, whose rows are read and
// written in 16-byte vectors: arrays that do not start on a 16-byte boundary
// (a view at an odd offset) take the whole-row kernels' scalar accesses
static bool band_path(int L, const void *a, const void *b = nullptr) {
    const bool aligned = ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0);
    return (L == 64 || L == 128 || L == 256) && aligned;
}
template <class R>
static int band_partials(int L) { return L / BandRows<R>::value; }

template <class R>
static void terms_band(void *stream, const R *th, int R_, int L, const double *twist, double *partials) {
    const dim3 grid((unsigned)(L / BandRows<R>::value), (unsigned)R_);
    if (L == 64)
        hipLaunchKernelGGL((xy_terms_band_kernel<R, 1>), grid, dim3(64), 0, as_stream(stream), th, twist, partials);
    else if (L == 128)
        hipLaunchKernelGGL((xy_terms_band_kernel<R, 2>), grid, dim3(64), 0, as_stream(stream), th, twist, partials);
    else
        hipLaunchKernelGGL((xy_terms_band_kernel<R, 4>), grid, dim3(64), 0, as_stream(stream), th, twist, partials);
}
template <class R>
static void grad_band(void *stream, const R *th, int R_, int L, const double *twist, Coefs k, R *grad) {
    const dim3 grid((unsigned)(L / BandRows<R>::value), (unsigned)R_);
    if (L == 64)
        hipLaunchKernelGGL((xy_grad_band_kernel<R, 1>), grid, dim3(64), 0, as_stream(stream), th, twist, k, grad);
    else if (L == 128)
        hipLaunchKernelGGL((xy_grad_band_kernel<R, 2>), grid, dim3(64), 0, as_stream(stream), th, twist, k, grad);
    else
        hipLaunchKernelGGL((xy_grad_band_kernel<R, 4>), grid, dim3(64), 0, as_stream(stream), th, twist, k, grad);
}

// Whole-row kernels take L = 64, 128, ..., 1024 (threads = L)
static bool rows_path(int L) { return L >= 64 && L <= 1024 && L % 64 == 0 && (L & (L - 1)) == 0; }

// Partials per replica of the whole-row terms kernel: one per wavefront
static int terms_rows_partials(int L) { return (L / kSegRows) * (L / 64); }

template <class R, int NT>
static void launch_terms_rows(void *stream, const R *th, int R_, const double *twist, double *partials) {
    hipLaunchKernelGGL((xy_terms_rows_kernel<R, NT>), dim3((unsigned)(NT / kSegRows), (unsigned)R_), dim3(NT), 0,
                       as_stream(stream), th, twist, partials);
}
template <class R>
static void terms_rows(void *stream, const R *th, int R_, int L, const double *twist, double *partials) {
    switch (L) {
    case 64: launch_terms_rows<R, 64>(stream, th, R_, twist, partials); break;
    case 128: launch_terms_rows<R, 128>(stream, th, R_, twist, partials); break;
    case 256: launch_terms_rows<R, 256>(stream, th, R_, twist, partials); break;
    case 512: launch_terms_rows<R, 512>(stream, th, R_, twist, partials); break;
    default: launch_terms_rows<R, 1024>(stream, th, R_, twist, partials); break;
    }
}
template <class R, int NT>
static void launch_grad_rows(void *stream, const R *th, int R_, const double *twist, Coefs k, R *grad) {
    hipLaunchKernelGGL((xy_grad_rows_kernel<R, NT>), dim3((unsigned)(NT / kSegRows), (unsigned)R_), dim3(NT), 0,
                       as_stream(stream), th, twist, k, grad);
}
template <class R>
static void grad_rows(void *stream, const R *th, int R_, int L, const double *twist, Coefs k, R *grad) {
    switch (L) {
    case 64: launch_grad_rows<R, 64>(stream, th, R_, twist, k, grad); break;
    case 128: launch_grad_rows<R, 128>(stream, th, R_, twist, k, grad); break;
    case 256: launch_grad_rows<R, 256>(stream, th, R_, twist, k, grad); break;
    case 512: launch_grad_rows<R, 512>(stream, th, R_, twist, k, grad); break;
    default: launch_grad_rows<R, 1024>(stream, th, R_, twist, k, grad); break;
    }
}

static int tiles_of(int L) { return ((L + kTileRows - 1) / kTileRows) * ((L + kTileCols - 1) / kTileCols); }
static dim3 site_grid(int L, int R_) { return dim3((unsigned)tiles_of(L), (unsigned)R_); }

template <class R, int COMP>
static int terms_impl(void *stream, const void *state, int R_, int L, const double *twist, double *out) {
    MC_REQUIRE(L >= 2, "invalid lattice size %d", L);
    MC_REQUIRE(R_ >= 1, "invalid replica count %d", R_);
    const bool band = COMP == 1 && band_path(L, state);
    const bool rows = COMP == 1 && !band && rows_path(L);
    const int ntiles = tiles_of(L);
    const int nblk = band ? band_partials<R>(L) : rows ? terms_rows_partials(L) : (ntiles + kTermsTiles - 1) / kTermsTiles;
    void *base = nullptr;
    int rc = reduction_scratch(stream, (size_t)R_ * nblk * 3 * sizeof(double), &base);
    if (rc) return rc;
    double *partials = reinterpret_cast<double *>(base);
    if (band)
        terms_band<R>(stream, (const R *)state, R_, L, twist, partials);
    else if (rows)
        terms_rows<R>(stream, (const R *)state, R_, L, twist, partials);
    else
        hipLaunchKernelGGL((terms_tile_kernel<R, COMP>), dim3((unsigned)nblk, (unsigned)R_), dim3(kTileThreads), 0,
                           as_stream(stream), (const R *)state, L, twist, ntiles, nblk, partials);
    MC_CHECK_LAUNCH();
    hipLaunchKernelGGL((reduce_finish_kernel<3, false, TermsOut>), dim3(R_), dim3(64), 0, as_stream(stream),
                       TermsOut{}, R_, nblk, nullptr, nullptr, out, partials);
    MC_CHECK_LAUNCH();
    return MC_OK;
}

}  // namespace mc

using namespace mc;

extern "C" {

int mc_xy_twist_terms(void *stream, int dtype, const void *theta, int R, int L, const double *twist,
                      double *out) {
    if (dtype == MC_F32) return terms_impl<float, 1>(stream, theta, R, L, twist, out);
    if (dtype == MC_F64) return terms_impl<double, 1>(stream, theta, R, L, twist, out);
    set_error(MC_ERR_TYPE, "unsupported XY dtype code %d", dtype);
    return MC_ERR_TYPE;
}

int mc_kawamura_twist_terms(void *stream, int dtype, const void *state, int R, int L, const double *twist,
                            double *out) {
    if (dtype == MC_F32) return terms_impl<float, 5>(stream, state, R, L, twist, out);
    if (dtype == MC_F64) return terms_impl<double, 5>(stream, state, R, L, twist, out);
    set_error(MC_ERR_TYPE, "unsupported Kawamura dtype code %d", dtype);
    return MC_ERR_TYPE;
}

int mc_xy_twist_grad(void *stream, int dtype, const void *theta, int R, int L, const double *twist,
                     const double *alpha, const double *beta, const double *gamma, void *grad) {
    MC_REQUIRE(R >= 1 && L >= 2, "invalid shape R=%d L=%d", R, L);
    const Coefs k{alpha, beta, gamma};
    if (dtype == MC_F32 && band_path(L, theta, grad))
        grad_band<float>(stream, (const float *)theta, R, L, twist, k, (float *)grad);
    else if (dtype == MC_F64 && band_path(L, theta, grad))
        grad_band<double>(stream, (const double *)theta, R, L, twist, k, (double *)grad);
    else if (dtype == MC_F32 && rows_path(L))
        grad_rows<float>(stream, (const float *)theta, R, L, twist, k, (float *)grad);
    else if (dtype == MC_F64 && rows_path(L))
        grad_rows<double>(stream, (const double *)theta, R, L, twist, k, (double *)grad);
    else if (dtype == MC_F32)
        hipLaunchKernelGGL(xy_grad_kernel<float>, site_grid(L, R), dim3(256), 0, as_stream(stream),
                           (const float *)theta, L, twist, k, (float *)grad);
    else if (dtype == MC_F64)
        hipLaunchKernelGGL(xy_grad_kernel<double>, site_grid(L, R), dim3(256), 0, as_stream(stream),
                           (const double *)theta, L, twist, k, (double *)grad);
    else {
        set_error(MC_ERR_TYPE, "unsupported XY dtype code %d", dtype);
        return MC_ERR_TYPE;
    }
    MC_CHECK_LAUNCH();
    return MC_OK;
}

int mc_xy_twist_hvp(void *stream, int dtype, const void *theta, int R, int L, const double *twist,
                    const double *alpha, const double *beta, const double *gamma, const void *u, void *out) {
    MC_REQUIRE(R >= 1 && L >= 2, "invalid shape R=%d L=%d", R, L);
    const Coefs k{alpha, beta, gamma};
    if (dtype == MC_F32)
        hipLaunchKernelGGL(xy_hvp_kernel<float>, site_grid(L, R), dim3(256), 0, as_stream(stream),
                           (const float *)theta, L, twist, k, (const float *)u, (float *)out);
    else if (dtype == MC_F64)
        hipLaunchKernelGGL(xy_hvp_kernel<double>, site_grid(L, R), dim3(256), 0, as_stream(stream),
                           (const double *)theta, L, twist, k, (const double *)u, (double *)out);
    else {
        set_error(MC_ERR_TYPE, "unsupported XY dtype code %d", dtype);
        return MC_ERR_TYPE;
    }
    MC_CHECK_LAUNCH();
    return MC_OK;
}

int mc_kawamura_twist_grad(void *stream, int dtype, const void *state, int R, int L, const double *twist,
                           const double *alpha, const double *beta, const double *gamma, void *grad) {
    MC_REQUIRE(R >= 1 && L >= 2, "invalid shape R=%d L=%d", R, L);
    const Coefs k{alpha, beta, gamma};
    if (dtype == MC_F32)
        hipLaunchKernelGGL(kawamura_grad_kernel<float>, site_grid(L, R), dim3(256), 0, as_stream(stream),
                           (const float *)state, L, twist, k, (float *)grad);
    else if (dtype == MC_F64)
        hipLaunchKernelGGL(kawamura_grad_kernel<double>, site_grid(L, R), dim3(256), 0, as_stream(stream),
                           (const double *)state, L, twist, k, (double *)grad);
    else {
        set_error(MC_ERR_TYPE, "unsupported Kawamura dtype code %d", dtype);
        return MC_ERR_TYPE;
    }
    MC_CHECK_LAUNCH();
    return MC_OK;
}

}  // extern "C"
