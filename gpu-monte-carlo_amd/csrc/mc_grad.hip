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
// One thread per site; its four neighbours come through L1/L2.  Per-site terms
// are evaluated in the state's precision, per-replica sums in float64.
#include "mc_reduce.h"

namespace mc {

template <class R>
__device__ __forceinline__ void sincos_r(R x, R &s, R &c);
// sine and cosine from one shared argument reduction
template <>
__device__ __forceinline__ void sincos_r<float>(float x, float &s, float &c) {
    ::sincosf(x, &s, &c);
}
template <>
__device__ __forceinline__ void sincos_r<double>(double x, double &s, double &c) {
    ::sincos(x, &s, &c);
}

__device__ __forceinline__ double twist_of(const double *twist, int r) { return twist ? twist[r] : 0.0; }

// ------------------------------------------------------------------ terms

template <class R, int COMP>
struct TwistAcc {
    const R *st;
    const double *twist;
    int L;
    template <bool ENERGY_ONLY, int CA>
    __device__ __forceinline__ void site(int r, int i, int j, double (&v)[CA], bool) const {
        const R *b = st + (size_t)r * L * L * COMP;
        const R *me = b + ((size_t)i * L + j) * COMP;
        const R *nx = b + ((size_t)wrap_dn(i, L) * L + j) * COMP;
        const R *ny = b + ((size_t)i * L + wrap_dn(j, L)) * COMP;
        const R dl = (R)twist_of(twist, r);
        R wx = 1, wy = 1;
        if (COMP == 5) {
            wx = ((R)1 + me[1] * nx[1]) * ((me[2] * nx[2] + me[3] * nx[3]) + me[4] * nx[4]);
            wy = ((R)1 + me[1] * ny[1]) * ((me[2] * ny[2] + me[3] * ny[3]) + me[4] * ny[4]);
        }
        R sx, cx, sy, cy;
        sincos_r<R>((me[0] - nx[0]) - dl, sx, cx);
        sincos_r<R>(me[0] - ny[0], sy, cy);
        v[0] += (double)(wx * cx);
        v[1] += (double)(wy * cy);
        v[2] += (double)(wx * sx);
    }
    template <int CA>
    __device__ __forceinline__ void finalize(const double (&t)[CA], double, double *out, int R_,
                                             int dest) const {
#pragma unroll
        for (int c = 0; c < CA; c++) out[(size_t)c * R_ + dest] = t[c];
    }
};

// ------------------------------------------------------------- gradient

// Per-replica coefficients; NULL means 0.
struct Coefs {
    const double *alpha, *beta, *gamma;
    __device__ __forceinline__ void at(int r, double &a, double &b, double &g) const {
        a = alpha ? alpha[r] : 0.0;
        b = beta ? beta[r] : 0.0;
        g = gamma ? gamma[r] : 0.0;
    }
};

template <class R>
__global__ __launch_bounds__(256) void xy_grad_kernel(const R *__restrict__ th, int L, const double *twist,
                                                      Coefs k, R *__restrict__ grad) {
    const int r = blockIdx.y;
    const int site = blockIdx.x * 256 + threadIdx.x;
    if (site >= L * L) return;
    const int i = site / L, j = site - (site / L) * L;
    const R *b = th + (size_t)r * L * L;
    const R t = b[site];
    const R tdn = b[wrap_dn(i, L) * L + j], tup = b[wrap_up(i, L) * L + j];
    const R trt = b[i * L + wrap_dn(j, L)], tlf = b[i * L + wrap_up(j, L)];
    double al, be, ga;
    k.at(r, al, be, ga);
    const R dl = (R)twist_of(twist, r);
    R s1, c1, s2, c2, s3, c3, s4, c4;
    sincos_r<R>((t - tdn) - dl, s1, c1);  // bond (i,j)-(i+1,j), site first
    sincos_r<R>((tup - t) - dl, s2, c2);  // bond (i-1,j)-(i,j), site second
    sincos_r<R>(t - trt, s3, c3);         // bond (i,j)-(i,j+1)
    sincos_r<R>(tlf - t, s4, c4);         // bond (i,j-1)-(i,j)
    // dCx/dt = -s1 + s2, dCy/dt = -s3 + s4, dSx/dt = c1 - c2
    const R g = (R)al * (s2 - s1) + (R)be * (s4 - s3) + (R)ga * (c1 - c2);
    grad[(size_t)r * L * L + site] = g;
}

template <class R>
__global__ __launch_bounds__(256) void xy_hvp_kernel(const R *__restrict__ th, int L, const double *twist,
                                                     Coefs k, const R *__restrict__ u, R *__restrict__ out) {
    const int r = blockIdx.y;
    const int site = blockIdx.x * 256 + threadIdx.x;
    if (site >= L * L) return;
    const int i = site / L, j = site - (site / L) * L;
    const size_t base = (size_t)r * L * L;
    const R *b = th + base;
    const R *v = u + base;
    const int kdn = wrap_dn(i, L) * L + j, kup = wrap_up(i, L) * L + j;
    const int krt = i * L + wrap_dn(j, L), klf = i * L + wrap_up(j, L);
    const R t = b[site];
    double al, be, ga;
    k.at(r, al, be, ga);
    const R dl = (R)twist_of(twist, r);
    R s1, c1, s2, c2, s3, c3, s4, c4;
    sincos_r<R>((t - b[kdn]) - dl, s1, c1);
    sincos_r<R>((b[kup] - t) - dl, s2, c2);
    sincos_r<R>(t - b[krt], s3, c3);
    sincos_r<R>(b[klf] - t, s4, c4);
    const R du_dn = v[site] - v[kdn], du_up = v[site] - v[kup];
    const R du_rt = v[site] - v[krt], du_lf = v[site] - v[klf];
    // Hess(Cx) u = -c1 du_dn - c2 du_up ; Hess(Cy) u = -c3 du_rt - c4 du_lf ;
    // Hess(Sx) u = -s1 du_dn - s2 du_up
    const R hx = -(c1 * du_dn) - c2 * du_up;
    const R hy = -(c3 * du_rt) - c4 * du_lf;
    const R hs = -(s1 * du_dn) - s2 * du_up;
    out[base + site] = (R)al * hx + (R)be * hy + (R)ga * hs;
}

// Kawamura: state (theta, tau, dx, dy, dz) per site.  For a bond term
// f(a) w with w = (1 + tau_p tau_q)(d_p . d_q), a = theta_p - theta_q - Delta_b:
//   d/dtheta_p = w f'(a), d/dtheta_q = -w f'(a),
//   d/dtau_p = tau_q (d_p . d_q) f(a), d/dd_p = (1 + tau_p tau_q) f(a) d_q.
template <class R>
__global__ __launch_bounds__(256) void kawamura_grad_kernel(const R *__restrict__ st, int L, const double *twist,
                                                            Coefs k, R *__restrict__ grad) {
    const int r = blockIdx.y;
    const int site = blockIdx.x * 256 + threadIdx.x;
    if (site >= L * L) return;
    const int i = site / L, j = site - (site / L) * L;
    const R *b = st + (size_t)r * L * L * 5;
    const R *me = b + (size_t)site * 5;
    double al, be, ga;
    k.at(r, al, be, ga);
    const R dl = (R)twist_of(twist, r);
    const R t = me[0], tau = me[1], dx = me[2], dy = me[3], dz = me[4];
    R g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0;
    // (neighbour, is_x_bond, site_is_first)
    const int nbr[4] = {wrap_dn(i, L) * L + j, wrap_up(i, L) * L + j, i * L + wrap_dn(j, L), i * L + wrap_up(j, L)};
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const R *o = b + (size_t)nbr[q] * 5;
        const bool xb = q < 2, first = (q & 1) == 0;
        const R a = first ? (t - o[0]) - (xb ? dl : (R)0) : (o[0] - t) - (xb ? dl : (R)0);
        R s, c;
        sincos_r<R>(a, s, c);
        const R dot = (dx * o[2] + dy * o[3]) + dz * o[4];
        const R tt = (R)1 + tau * o[1];
        const R w = tt * dot;
        // f = cos with coefficient (alpha | beta), f = sin with gamma (x bonds only)
        const R kc = (R)(xb ? al : be), ks = xb ? (R)ga : (R)0;
        const R f = kc * c + ks * s;       // coefficient-weighted f(a)
        const R fp = ks * c - kc * s;      // coefficient-weighted f'(a)
        g0 += first ? w * fp : -(w * fp);
        g1 += o[1] * dot * f;
        g2 += tt * f * o[2];
        g3 += tt * f * o[3];
        g4 += tt * f * o[4];
    }
    R *out = grad + ((size_t)r * L * L + site) * 5;
    out[0] = g0;
    out[1] = g1;
    out[2] = g2;
    out[3] = g3;
    out[4] = g4;
}

template <class R, int COMP>
static int terms_impl(void *stream, const void *state, int R_, int L, const double *twist, double *out) {
    MC_REQUIRE(L >= 2, "invalid lattice size %d", L);
    return launch_reduce<3, false>(stream, TwistAcc<R, COMP>{(const R *)state, twist, L}, R_, L, nullptr,
                                   nullptr, out);
}

static dim3 site_grid(int L, int R_) { return dim3((unsigned)(((long)L * L + 255) / 256), (unsigned)R_); }

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
    if (dtype == MC_F32)
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
