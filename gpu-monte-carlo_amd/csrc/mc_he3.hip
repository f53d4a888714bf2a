// He3 thin-film Ginzburg-Landau model: 3x2 complex order parameter per site
// (12 reals, re/im interleaved).  Cartesian Metropolis, conditional-Gaussian
// MH, and the ten measurement channels.
#include "mc_common.h"
#include "mc_reduce.h"

namespace mc {

template <class R> struct He3Coef {
    R b[5];
    R K, g_d;
};

// pointer to component (mu, a) of site (i, j) inside one replica block
template <class R>
__device__ __forceinline__ Cpx<R> ld(const R *__restrict__ blk, int L, int i, int j, int mu, int a) {
    const R *p = blk + 12 * (i * L + j) + 2 * (2 * mu + a);
    return Cpx<R>{p[0], p[1]};
}

// dipole invariant with spin rows x, y and orbital columns x, y
template <class R>
__device__ __forceinline__ R dipole_invariant(const Cpx<R> *psi, R n) {
    const Cpx<R> a = psi[0], b = psi[1], c = psi[2], d = psi[3];
    return (((norm2(a + d) + norm2(a)) + norm2(d)) + R(2) * (b * conj(c)).re) - R(2.0 / 3.0) * n;
}

// -2n + sum beta_k I_k + g_d * dipole (psi[mu*2 + a])
template <class R>
__device__ __forceinline__ R potential(const Cpx<R> *psi, const He3Coef<R> &c) {
    R nx = norm2(psi[0]), ny = norm2(psi[1]);
    Cpx<R> xx = psi[0] * psi[0], yy = psi[1] * psi[1];
    Cpx<R> xy = psi[0] * psi[1];
    Cpx<R> overlap = conj(psi[0]) * psi[1];
#pragma unroll
    for (int mu = 1; mu < 3; mu++) {
        const Cpx<R> x = psi[2 * mu], y = psi[2 * mu + 1];
        nx += norm2(x);
        ny += norm2(y);
        xx = xx + x * x;
        yy = yy + y * y;
        xy = xy + x * y;
        overlap = overlap + conj(x) * y;
    }
    const R n = nx + ny, two = R(2);
    const R i1 = norm2(xx + yy);
    const R i2 = n * n;
    const R i3 = (nx * nx + ny * ny) + two * (overlap * overlap).re;
    const R i4 = (nx * nx + ny * ny) + two * norm2(overlap);
    const R i5 = (norm2(xx) + norm2(yy)) + two * norm2(xy);
    const R dip = dipole_invariant(psi, n);
    return (((((-two * n + c.b[0] * i1) + c.b[1] * i2) + c.b[2] * i3) + c.b[3] * i4) + c.b[4] * i5) +
           c.g_d * dip;
}

// three forward-difference gradient contractions of one spin row
template <class R>
__device__ __forceinline__ R gradient(Cpx<R> dx_x, Cpx<R> dx_y, Cpx<R> dy_x, Cpx<R> dy_y) {
    return ((R(3) * (norm2(dx_x) + norm2(dy_y)) + norm2(dx_y)) + norm2(dy_x)) +
           R(2) * ((conj(dx_x) * dy_y).re + (conj(dx_y) * dy_x).re);
}

// Change of the three gradient cells touched by site (i, j) for spin row
// mu: the site's old (x, y) and proposed values (ex, ey are their
// differences), and the neighbours right (ip, j), up (i, jp), left (im, j),
// down (i, jm), left-up (im, jp) and right-down (ip, jm).
template <class R>
__device__ __forceinline__ R gradient_cells(Cpx<R> x, Cpx<R> y, Cpx<R> ex, Cpx<R> ey, Cpx<R> rx, Cpx<R> ry,
                                            Cpx<R> ux, Cpx<R> uy, Cpx<R> lx, Cpx<R> ly, Cpx<R> dx, Cpx<R> dy,
                                            Cpx<R> lux, Cpx<R> luy, Cpx<R> drx, Cpx<R> dry) {
    const R three = R(3), two = R(2), eight = R(8), four = R(4);
    const Cpx<R> a = rx - x, b = ry - y, c = ux - x, d = uy - y;
    const Cpx<R> al = x - lx, bl = y - ly, cl = lux - lx, dl = luy - ly;
    const Cpx<R> ad = drx - dx, bd = dry - dy, cd = x - dx, dd = y - dy;
    const Cpx<R> t1 = ((scale(three, a) + b) + c) + d;
    const Cpx<R> fx = (neg(t1) + (scale(three, al) + dl)) + (cd + bd);
    const Cpx<R> t2 = ((a + b) + c) + scale(three, d);
    const Cpx<R> fy = (neg(t2) + (bl + cl)) + (scale(three, dd) + ad);
    const R linear = two * ((conj(ex) * fx).re + (conj(ey) * fy).re);
    const R quadratic = eight * (norm2(ex) + norm2(ey)) + four * (conj(ex) * ey).re;
    return linear + quadratic;
}

// exact change of the three gradient cells touched by site (i, j)
template <class R>
__device__ __forceinline__ R gradient_difference(const R *__restrict__ blk, int L, int i, int j,
                                                 const Cpx<R> *prop) {
    const int ip = wrap_dn(i, L), im = wrap_up(i, L), jp = wrap_dn(j, L), jm = wrap_up(j, L);
    R delta_g = R(0);
#pragma unroll
    for (int mu = 0; mu < 3; mu++) {
        const Cpx<R> x = ld(blk, L, i, j, mu, 0), y = ld(blk, L, i, j, mu, 1);
        delta_g += gradient_cells(x, y, prop[2 * mu] - x, prop[2 * mu + 1] - y, ld(blk, L, ip, j, mu, 0),
                                  ld(blk, L, ip, j, mu, 1), ld(blk, L, i, jp, mu, 0), ld(blk, L, i, jp, mu, 1),
                                  ld(blk, L, im, j, mu, 0), ld(blk, L, im, j, mu, 1), ld(blk, L, i, jm, mu, 0),
                                  ld(blk, L, i, jm, mu, 1), ld(blk, L, im, jp, mu, 0), ld(blk, L, im, jp, mu, 1),
                                  ld(blk, L, ip, jm, mu, 0), ld(blk, L, ip, jm, mu, 1));
    }
    return delta_g;
}

// The site and the six neighbours its gradient cells read, in the order
// self, right, up, left, down, left-up, right-down.
constexpr int kStencil = 7;

// the same from a stencil loaded beforehand (load_stencil)
template <class R>
__device__ __forceinline__ R gradient_difference(const Cpx<R> (&s)[kStencil][6], const Cpx<R> *prop) {
    R delta_g = R(0);
#pragma unroll
    for (int mu = 0; mu < 3; mu++) {
        const int a = 2 * mu, b = 2 * mu + 1;
        delta_g += gradient_cells(s[0][a], s[0][b], prop[a] - s[0][a], prop[b] - s[0][b], s[1][a], s[1][b],
                                  s[2][a], s[2][b], s[3][a], s[3][b], s[4][a], s[4][b], s[5][a], s[5][b],
                                  s[6][a], s[6][b]);
    }
    return delta_g;
}

template <class R>
__device__ __forceinline__ void load_site(const R *__restrict__ blk, int L, int i, int j, Cpx<R> *psi) {
#pragma unroll
    for (int m = 0; m < 6; m++) psi[m] = ld(blk, L, i, j, m >> 1, m & 1);
}

template <class R>
__device__ __forceinline__ void load_stencil(const R *__restrict__ blk, int L, int i, int j,
                                             Cpx<R> (&s)[kStencil][6]) {
    const int ip = wrap_dn(i, L), im = wrap_up(i, L), jp = wrap_dn(j, L), jm = wrap_up(j, L);
    const int si[kStencil] = {i, ip, i, im, i, im, ip}, sj[kStencil] = {j, j, jp, j, jm, jp, jm};
#pragma unroll
    for (int n = 0; n < kStencil; n++) load_site(blk, L, si[n], sj[n], s[n]);
}

template <class R>
__device__ __forceinline__ void store_site(R *__restrict__ blk, int L, int i, int j, const Cpx<R> *psi) {
    R *p = blk + 12 * (i * L + j);
#pragma unroll
    for (int m = 0; m < 6; m++) {
        p[2 * m] = psi[m].re;
        p[2 * m + 1] = psi[m].im;
    }
}

// One Cartesian Metropolis visit of a site (models/he3/update.py:186-227)
// from its stencil (load_stencil's order) and its stream: the box proposal
// (12 draws, re and im per component) into psi, then one acceptance draw;
// returns the decision.
template <class R>
__device__ __forceinline__ bool he3_visit(const Cpx<R> (&s)[kStencil][6], Rng &g, R beta, const He3Coef<R> &c,
                                          R delta, Cpx<R> (&psi)[6]) {
#pragma unroll
    for (int m = 0; m < 6; m++) psi[m] = s[0][m];
    const R old_potential = potential(psi, c);
#pragma unroll
    for (int m = 0; m < 6; m++) {
        const R re = delta * (R(2) * g.template uniform<R>() - R(1));
        const R im = delta * (R(2) * g.template uniform<R>() - R(1));
        psi[m] = Cpx<R>{psi[m].re + re, psi[m].im + im};
    }
    const R dpot = potential(psi, c) - old_potential;
    const R dgrad = gradient_difference(s, psi);
    const R diff = dpot + c.K * dgrad;
    const R u = g.template uniform<R>();
    return metropolis_accept(diff, u, beta);
}

template <class R>
__global__ __launch_bounds__(256) void he3_color_kernel(R *__restrict__ st, uint64_t *__restrict__ rng,
                                                        const R *__restrict__ betas, int L, int color,
                                                        He3Coef<R> c, R delta) {
    const int h = L >> 1;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= h * h) return;
    const int r = blockIdx.y;
    const int row = t / h, col = t - (t / h) * h;
    const int i = 2 * row + (color >> 1), j = 2 * col + (color & 1);
    R *blk = st + (size_t)r * L * L * 12;
    const size_t k = (size_t)r * L * L + (size_t)i * L + j;
    // Every load is issued before any arithmetic, the stream first (the
    // draws need only it) and the stencil after, so the site waits for one
    // memory round trip rather than one per neighbour and component; at the
    // small lattices of the reference's benchmark (L = 32, 100 replicas: less
    // than one wave per SIMD) nothing else hides those waits.
    Rng g = load_rng(rng, k);
    const R beta = betas[r];
    Cpx<R> s[kStencil][6];
    load_stencil(blk, L, i, j, s);
    __builtin_amdgcn_sched_barrier(0);
    Cpx<R> psi[6];
    const bool accept = he3_visit(s, g, beta, c, delta, psi);
    store_rng(rng, k, g);
    if (accept) store_site(blk, L, i, j, psi);
}

// Half a sweep of the Cartesian update in one launch: colours 2 half then
// 2 half + 1 of the rows of parity `half`, a workgroup owning whole rows.  A
// site's stencil reaches rows i - 1 and i + 1 only, which have the other
// parity and do not change in this launch, and its own row's sites of the
// other colour, which the workgroup's first pass updates.  So two launches
// per sweep leave what four colour launches leave, bit for bit: the visit is
// the colour kernel's.  Thread (p, q) of the block of rows from i0 visits
// row i = i0 + half + 2 p, column 2 q (first pass) and 2 q + 1 (second).
// The second site's stencil is
//   self (i, 2q+1), left (i-1, 2q+1) and right-down (i+1, 2q): in the
//       first site's stencil (its up, left-up and right neighbours);
//   right (i+1, 2q+1) and left-up (i-1, 2q+2): loaded with the first
//       stencil, in the same memory round trip;
//   down (i, 2q): the thread's own first site after its visit;
//   up (i, 2q+2): thread (p, q+1 mod L/2)'s first site after its visit,
//       through LDS behind the workgroup barrier;
// so the second pass waits for no memory round trip.  float64 keeps the
// first form, the second stencil loaded after the barrier (a workgroup-scope
// release and acquire orders the first pass's stores before those loads):
// holding three of the first stencil's sites across the first visit spills
// its registers.
template <class R>
__global__ __launch_bounds__(512) void he3_rows_kernel(R *__restrict__ st, uint64_t *__restrict__ rng,
                                                       const R *__restrict__ betas, int L, int B, int half,
                                                       He3Coef<R> c, R delta) {
    const int r = blockIdx.y, h = L >> 1;
    const int p = threadIdx.x / h, q = threadIdx.x - (threadIdx.x / h) * h;
    const int i = blockIdx.x * B + half + 2 * p;
    R *blk = st + (size_t)r * L * L * 12;
    const size_t k0 = (size_t)r * L * L + (size_t)i * L + 2 * q;
    const R beta = betas[r];
    if constexpr (sizeof(R) == 8) {
#pragma unroll
        for (int cc = 0; cc < 2; cc++) {
            const int j = 2 * q + cc;
            Rng g = load_rng(rng, k0 + cc);
            Cpx<R> s[kStencil][6];
            load_stencil(blk, L, i, j, s);
            __builtin_amdgcn_sched_barrier(0);
            Cpx<R> psi[6];
            const bool accept = he3_visit(s, g, beta, c, delta, psi);
            store_rng(rng, k0 + cc, g);
            if (accept) store_site(blk, L, i, j, psi);
            if (cc == 0) __syncthreads();
        }
    } else {
        __shared__ Cpx<R> first[512][6];
        const int ip = wrap_dn(i, L), im = wrap_up(i, L), jb = 2 * q + 1, jn = wrap_dn(jb, L);
        Rng ga = load_rng(rng, k0), gb = load_rng(rng, k0 + 1);
        Cpx<R> s[kStencil][6], right_b[6], left_up_b[6];
        load_stencil(blk, L, i, 2 * q, s);
        load_site(blk, L, ip, jb, right_b);
        load_site(blk, L, im, jn, left_up_b);
        __builtin_amdgcn_sched_barrier(0);
        // the first site's value after its visit, in LDS: the old value now,
        // the proposal if accepted (a select of the two would put both arrays
        // in scratch memory)
#pragma unroll
        for (int m = 0; m < 6; m++) first[threadIdx.x][m] = s[0][m];
        Cpx<R> psi[6];
        const bool accept_a = he3_visit(s, ga, beta, c, delta, psi);
        // every load has long arrived; waiting for them here keeps the second
        // visit from waiting on this visit's stores (vmcnt counts both, in
        // order)
        __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
        store_rng(rng, k0, ga);
        if (accept_a) {
            store_site(blk, L, i, 2 * q, psi);
#pragma unroll
            for (int m = 0; m < 6; m++) first[threadIdx.x][m] = psi[m];
        }
        __syncthreads();
        const int up_thread = p * h + (q + 1 == h ? 0 : q + 1);
        Cpx<R> t[kStencil][6];
#pragma unroll
        for (int m = 0; m < 6; m++) {
            t[0][m] = s[2][m];
            t[1][m] = right_b[m];
            t[2][m] = first[up_thread][m];
            t[3][m] = s[5][m];
            t[4][m] = first[threadIdx.x][m];
            t[5][m] = left_up_b[m];
            t[6][m] = s[1][m];
        }
        const bool accept_b = he3_visit(t, gb, beta, c, delta, psi);
        store_rng(rng, k0 + 1, gb);
        if (accept_b) store_site(blk, L, i, jb, psi);
    }
}

// Rows per workgroup of the He3 row kernel: among the even divisors B of L
// whose workgroup of (B/2) * (L/2) threads stays within 512, the largest that
// still gives every compute unit a workgroup (R * L / B >= cus), else the
// smallest; 0 when none fits (the colour kernels run instead).  At L = 64,
// R = 100: B = 16, 400 workgroups of 256 threads over 256 CUs, rather than
// 200 workgroups of 512 threads on 200 of them.
static int he3_rows_per_block(int L, int R, int cus) {
    int spread = 0, smallest = 0;
    for (int B = 2; B <= L; B += 2) {
        if (L % B || (B / 2) * (L / 2) > 512) continue;
        if (!smallest) smallest = B;
        if ((long long)R * (L / B) >= cus) spread = B;
    }
    return spread ? spread : smallest;
}

// conditional field of the fixed neighbours (gaussian_fast.py:26-40)
template <class R>
__device__ __forceinline__ void neighbor_field(const R *__restrict__ blk, int L, int i, int j, Cpx<R> *field) {
    const int ip = wrap_dn(i, L), im = wrap_up(i, L), jp = wrap_dn(j, L), jm = wrap_up(j, L);
    const R three = R(3);
#pragma unroll
    for (int mu = 0; mu < 3; mu++) {
        const Cpx<R> rx = ld(blk, L, ip, j, mu, 0), ry = ld(blk, L, ip, j, mu, 1);
        const Cpx<R> ux = ld(blk, L, i, jp, mu, 0), uy = ld(blk, L, i, jp, mu, 1);
        const Cpx<R> lx = ld(blk, L, im, j, mu, 0), ly = ld(blk, L, im, j, mu, 1);
        const Cpx<R> dx = ld(blk, L, i, jm, mu, 0), dy = ld(blk, L, i, jm, mu, 1);
        const Cpx<R> lux = ld(blk, L, im, jp, mu, 0), luy = ld(blk, L, im, jp, mu, 1);
        const Cpx<R> drx = ld(blk, L, ip, jm, mu, 0), dry = ld(blk, L, ip, jm, mu, 1);
        Cpx<R> f = scale(-three, rx + lx);
        f = f - ux; f = f - uy; f = f - ry; f = f - ly; f = f - dx; f = f - dy; f = f + luy; f = f + dry;
        field[2 * mu] = f;
        Cpx<R> q = neg(rx);
        q = q - ry; q = q - ux; q = q - lx; q = q - ly; q = q - dx; q = q - scale(three, uy + dy);
        q = q + lux; q = q + drx;
        field[2 * mu + 1] = q;
    }
}

template <class R>
__device__ __forceinline__ Cpx<R> normal_pair(Rng &g) {
    R u = g.template uniform<R>();
    if (u <= R(0)) u = R(1.1102230246251565e-16);
    const R angle = R(6.283185307179586) * g.template uniform<R>();
    const R radius = rsqrt_(-R(2) * rlog(u));
    return Cpx<R>{radius * rcos(angle), radius * rsin(angle)};
}

template <class R>
__device__ __forceinline__ R gaussian_proposal(Cpx<R> *psi, const Cpx<R> *field, R beta, R K, R rho, R mass,
                                               Rng &g) {
    const R one = R(1), two = R(2);
    const R root = R(0.7071067811865475);
    const R plus = R(10) * K + mass, minus = R(6) * K + mass;
    const R vf = one - rho * rho;
    const R sigma_plus = rsqrt_(vf / (two * beta * plus));
    const R sigma_minus = rsqrt_(vf / (two * beta * minus));
    const R q_plus = vf / ((((R(8) * root) * root) * sigma_plus) * sigma_plus);
    const R q_minus = vf / ((((R(8) * root) * root) * sigma_minus) * sigma_minus);
    R log_ratio = R(0);
#pragma unroll
    for (int mu = 0; mu < 3; mu++) {
        const Cpx<R> x = psi[2 * mu], y = psi[2 * mu + 1];
        const Cpx<R> hp = scale(root, field[2 * mu] + field[2 * mu + 1]);
        const Cpx<R> hm = scale(root, field[2 * mu] - field[2 * mu + 1]);
        const Cpx<R> kp = scale(-K, hp), km = scale(-K, hm);
        const Cpx<R> mp = Cpx<R>{kp.re / plus, kp.im / plus};
        const Cpx<R> mm = Cpx<R>{km.re / minus, km.im / minus};
        const Cpx<R> mx = scale(root, mp + mm);
        const Cpx<R> my = scale(root, mp - mm);
        const Cpx<R> zp = scale(sigma_plus, normal_pair<R>(g));
        const Cpx<R> zm = scale(sigma_minus, normal_pair<R>(g));
        const Cpx<R> nx = (mx + scale(rho, x - mx)) + scale(root, zp + zm);
        const Cpx<R> ny = (my + scale(rho, y - my)) + scale(root, zp - zm);
        psi[2 * mu] = nx;
        psi[2 * mu + 1] = ny;
        const Cpx<R> old_plus = (x - mx) + (y - my);
        const Cpx<R> old_minus = (x - mx) - (y - my);
        const Cpx<R> new_plus = (nx - mx) + (ny - my);
        const Cpx<R> new_minus = (nx - mx) - (ny - my);
        log_ratio += q_plus * (norm2(new_plus) - norm2(old_plus));
        log_ratio += q_minus * (norm2(new_minus) - norm2(old_minus));
    }
    return log_ratio;
}

// Occupancy: the complex64 update needs 167 VGPRs unconstrained (three waves
// per SIMD); asking for four waves (128 VGPRs, a few spills) measured 24%
// faster on MI355X, five waves and a constrained complex128 update slower.
template <class R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 1))) void
he3_gaussian_color_kernel(R *__restrict__ st, uint64_t *__restrict__ rng,
                                                                 const R *__restrict__ betas, int L,
                                                                 int color, He3Coef<R> c, R rho, R mass,
                                                                 int64_t *__restrict__ counts, size_t N) {
    const int h = L >> 1;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= h * h) return;
    const int r = blockIdx.y;
    const int row = t / h, col = t - (t / h) * h;
    const int i = 2 * row + (color >> 1), j = 2 * col + (color & 1);
    R *blk = st + (size_t)r * L * L * 12;
    const size_t k = (size_t)r * L * L + (size_t)i * L + j;
    const R beta = betas[r];
    Rng g = load_rng(rng, k);
    Cpx<R> field[6], psi[6];
    neighbor_field(blk, L, i, j, field);
    load_site(blk, L, i, j, psi);
    const R old_potential = potential(psi, c);
    const R lq = gaussian_proposal(psi, field, beta, c.K, rho, mass, g);
    const R dh = (potential(psi, c) - old_potential) + c.K * gradient_difference(blk, L, i, j, psi);
    const R log_ratio = -beta * dh + lq;
    const R u = g.template uniform<R>();
    store_rng(rng, k, g);
    // nonfinite: counted and rejected; else log r >= 0 or u < exp(log r),
    // evaluated without short-circuit control flow (see metropolis_accept)
    const bool finite = isfinite(log_ratio);
    const R x = log_ratio > R(0) ? R(0) : log_ratio;
    const bool accept = finite & ((log_ratio >= R(0)) | (u < rexp(x)));
    // The per-site counters by adds that return nothing (one thread per site,
    // so the same counts as a read-modify-write, without waiting for a read
    // of each counter at the end of every thread).
    unsigned long long *cnt = reinterpret_cast<unsigned long long *>(counts);
    atomicAdd(cnt + k, 1ull);
    if (!finite) atomicAdd(cnt + 2 * N + k, 1ull);
    if (accept) {
        atomicAdd(cnt + N + k, 1ull);
        store_site(blk, L, i, j, psi);
    }
}

// Accumulated channels: energy, density, chirality, phase re/im, current_x,
// curvature_x, current_y, curvature_y, spin-z density.
template <class R>
struct He3Acc {
    const R *st;
    He3Coef<R> c;
    int L;
    template <bool ENERGY_ONLY, int CA>
    __device__ __forceinline__ void site(int r, int i, int j, double (&v)[CA], bool cached) const {
        const R *blk = st + (size_t)r * L * L * 12;
        Cpx<R> psi[6];
        load_site(blk, L, i, j, psi);
        if (!cached) v[0] += (double)potential(psi, c);
        const int ip = wrap_dn(i, L), jp = wrap_dn(j, L);
#pragma unroll
        for (int mu = 0; mu < 3; mu++) {
            const Cpx<R> x = psi[2 * mu], y = psi[2 * mu + 1];
            const Cpx<R> xp_x = ld(blk, L, ip, j, mu, 0), xp_y = ld(blk, L, ip, j, mu, 1);
            const Cpx<R> yp_x = ld(blk, L, i, jp, mu, 0), yp_y = ld(blk, L, i, jp, mu, 1);
            const Cpx<R> a = xp_x - x, b = xp_y - y, cc = yp_x - x, d = yp_y - y;
            if (!cached) v[0] += (double)(c.K * gradient(a, b, cc, d));
            if (ENERGY_ONLY) continue;
            const R row_density = norm2(x) + norm2(y);
            v[1] += (double)row_density;
            if (mu == 2) v[9] += (double)row_density;
            v[2] += 2.0 * (double)(conj(x) * y).im;
            const Cpx<R> phase = x * x - y * y;
            v[3] += (double)phase.re;
            v[4] += (double)phase.im;
            const double ar = a.re, ai = -a.im, dr = d.re, di = -d.im;
            const double t1r = 3.0 * ar + dr, t1i = 3.0 * ai + di;
            const Cpx<R> bcx = (conj(b) + conj(cc)) * xp_y;
            const double zxr = (t1r * (double)xp_x.re - t1i * (double)xp_x.im) + (double)bcx.re;
            const double zxi = (t1r * (double)xp_x.im + t1i * (double)xp_x.re) + (double)bcx.im;
            const Cpx<R> cby = (conj(cc) + conj(b)) * yp_x;
            const double t2r = 3.0 * dr + ar, t2i = 3.0 * di + ai;
            const double zyr = (double)cby.re + (t2r * (double)yp_y.re - t2i * (double)yp_y.im);
            const double zyi = (double)cby.im + (t2r * (double)yp_y.im + t2i * (double)yp_y.re);
            const double twoK = 2.0 * (double)c.K;
            v[5] += twoK * zxi;
            v[6] += twoK * ((3.0 * (double)norm2(xp_x) + (double)norm2(xp_y)) - zxr);
            v[7] += twoK * zyi;
            v[8] += twoK * (((double)norm2(yp_x) + 3.0 * (double)norm2(yp_y)) - zyr);
        }
    }
    template <int CA>
    __device__ __forceinline__ void finalize(const double (&t)[CA], double energy, double *out, int R_,
                                             int dest) const {
        out[dest] = energy;
#pragma unroll
        for (int ch = 1; ch < 9; ch++) out[(size_t)ch * R_ + dest] = t[ch];
        out[(size_t)9 * R_ + dest] = t[1] == 0.0 ? 0.0 : 0.5 * (3 * t[9] / t[1] - 1);
    }
};

template <class R>
static He3Coef<R> make_coef(const double *coef, double K, double g_d) {
    He3Coef<R> c;
    for (int k = 0; k < 5; k++) c.b[k] = (R)coef[k];
    c.K = (R)K;
    c.g_d = (R)g_d;
    return c;
}

template <class R>
static int he3_sweep(void *stream, void *state, uint64_t *rng, const void *betas, int R_, int L,
                     const He3Coef<R> &c, R delta) {
    static int cus[64] = {0};
    int dev = 0;
    MC_TRY_HIP(hipGetDevice(&dev));
    if (dev >= 0 && dev < 64 && cus[dev] == 0)
        MC_TRY_HIP(hipDeviceGetAttribute(&cus[dev], hipDeviceAttributeMultiprocessorCount, dev));
    const int B = he3_rows_per_block(L, R_, dev >= 0 && dev < 64 ? cus[dev] : 256);
    const int h = L / 2;
    if (B == 0) {
        dim3 grid((unsigned)((h * h + 255) / 256), (unsigned)R_);
        for (int color = 0; color < 4; color++) {
            hipLaunchKernelGGL(he3_color_kernel<R>, grid, dim3(256), 0, as_stream(stream), (R *)state, rng,
                               (const R *)betas, L, color, c, delta);
            MC_CHECK_LAUNCH();
        }
        return MC_OK;
    }
    const int threads = (B / 2) * h;
    for (int half = 0; half < 2; half++) {
        hipLaunchKernelGGL(he3_rows_kernel<R>, dim3((unsigned)(L / B), (unsigned)R_), dim3(threads), 0,
                           as_stream(stream), (R *)state, rng, (const R *)betas, L, B, half, c, delta);
        MC_CHECK_LAUNCH();
    }
    return MC_OK;
}

}  // namespace mc

using namespace mc;

#define HE3_CHECK_SHAPE()                                                      \
    MC_REQUIRE(L >= 2 && L % 2 == 0, "L must be even, got %d", L);             \
    MC_REQUIRE(R >= 1, "need at least one replica");                          \
    MC_REQUIRE(color >= 0 && color < 4, "color must be in 0..3, got %d", color); \
    MC_REQUIRE(coef != nullptr, "He3 coefficients are required")

extern "C" {

int mc_he3_update_color(void *stream, int dtype, void *state, uint64_t *rng, const void *betas, int R,
                        int L, int color, const double *coef, double K, double delta, double g_d) {
    HE3_CHECK_SHAPE();
    const int h = L / 2;
    dim3 grid((unsigned)((h * h + 255) / 256), (unsigned)R);
    if (dtype == MC_F32) {
        hipLaunchKernelGGL(he3_color_kernel<float>, grid, dim3(256), 0, as_stream(stream), (float *)state,
                           rng, (const float *)betas, L, color, make_coef<float>(coef, K, g_d), (float)delta);
    } else if (dtype == MC_F64) {
        hipLaunchKernelGGL(he3_color_kernel<double>, grid, dim3(256), 0, as_stream(stream), (double *)state,
                           rng, (const double *)betas, L, color, make_coef<double>(coef, K, g_d), delta);
    } else {
        set_error(MC_ERR_TYPE, "unsupported He3 dtype code %d", dtype);
        return MC_ERR_TYPE;
    }
    MC_CHECK_LAUNCH();
    return MC_OK;
}

int mc_he3_sweep(void *stream, int dtype, void *state, uint64_t *rng, const void *betas, int R, int L,
                 const double *coef, double K, double delta, double g_d) {
    MC_REQUIRE(L >= 2 && L % 2 == 0, "L must be even, got %d", L);
    MC_REQUIRE(R >= 1, "need at least one replica");
    MC_REQUIRE(coef != nullptr, "He3 coefficients are required");
    if (dtype == MC_F32)
        return he3_sweep<float>(stream, state, rng, betas, R, L, make_coef<float>(coef, K, g_d), (float)delta);
    if (dtype == MC_F64)
        return he3_sweep<double>(stream, state, rng, betas, R, L, make_coef<double>(coef, K, g_d), delta);
    set_error(MC_ERR_TYPE, "unsupported He3 dtype code %d", dtype);
    return MC_ERR_TYPE;
}

int mc_he3_gaussian_update_color(void *stream, int dtype, void *state, uint64_t *rng, const void *betas,
                                 int R, int L, int color, const double *coef, double K, double rho,
                                 double mass, int64_t *counts, double g_d) {
    HE3_CHECK_SHAPE();
    MC_REQUIRE(counts != nullptr, "Gaussian sampler needs a counts buffer");
    const int h = L / 2;
    const size_t N = (size_t)R * L * L;
    dim3 grid((unsigned)((h * h + 255) / 256), (unsigned)R);
    if (dtype == MC_F32) {
        hipLaunchKernelGGL(he3_gaussian_color_kernel<float>, grid, dim3(256), 0, as_stream(stream),
                           (float *)state, rng, (const float *)betas, L, color,
                           make_coef<float>(coef, K, g_d), (float)rho, (float)mass, counts, N);
    } else if (dtype == MC_F64) {
        hipLaunchKernelGGL(he3_gaussian_color_kernel<double>, grid, dim3(256), 0, as_stream(stream),
                           (double *)state, rng, (const double *)betas, L, color,
                           make_coef<double>(coef, K, g_d), rho, mass, counts, N);
    } else {
        set_error(MC_ERR_TYPE, "unsupported He3 dtype code %d", dtype);
        return MC_ERR_TYPE;
    }
    MC_CHECK_LAUNCH();
    return MC_OK;
}

int mc_he3_measure(void *stream, int dtype, const void *state, int R, int L, const double *coef,
                   double K, double g_d, const int32_t *temperature_indices,
                   const double *cached_energies, double *out) {
    MC_REQUIRE(coef != nullptr, "He3 coefficients are required");
    if (dtype == MC_F32)
        return launch_reduce<10, false>(stream,
                                        He3Acc<float>{(const float *)state, make_coef<float>(coef, K, g_d), L},
                                        R, L, temperature_indices, cached_energies, out);
    if (dtype == MC_F64)
        return launch_reduce<10, false>(
            stream, He3Acc<double>{(const double *)state, make_coef<double>(coef, K, g_d), L}, R, L,
            temperature_indices, cached_energies, out);
    set_error(MC_ERR_TYPE, "unsupported He3 dtype code %d", dtype);
    return MC_ERR_TYPE;
}

int mc_he3_energy(void *stream, int dtype, const void *state, int R, int L, const double *coef, double K,
                  double g_d, double *out) {
    MC_REQUIRE(coef != nullptr, "He3 coefficients are required");
    if (dtype == MC_F32)
        return launch_reduce<10, true>(stream,
                                       He3Acc<float>{(const float *)state, make_coef<float>(coef, K, g_d), L},
                                       R, L, nullptr, nullptr, out);
    if (dtype == MC_F64)
        return launch_reduce<10, true>(
            stream, He3Acc<double>{(const double *)state, make_coef<double>(coef, K, g_d), L}, R, L, nullptr,
            nullptr, out);
    set_error(MC_ERR_TYPE, "unsupported He3 dtype code %d", dtype);
    return MC_ERR_TYPE;
}

}  // extern "C"
