/*
 * CPU oracle for the batched float32 fast path (test infrastructure only).
 *
 * The fast XY update differs from the reference formula only in how it
 * evaluates dE:
 *   E_loc(x) = -J sum_k cos(x - theta_k) = -J (cos x * C + sin x * S),
 *   C = sum_k cos theta_k, S = sum_k sin theta_k,
 * with every site's (cos theta, sin theta) cached and refreshed on
 * acceptance, and with the float32 functions below in place of libm.  The
 * proposal, draws, wrap and acceptance rule are those of
 * models/xy/update.py:6-54.
 *
 * fast_sincosf / fast_expf are written from the mathematics: Cody-Waite
 * reduction by pi/2 (three-part split) or ln 2 (two-part split) with fused
 * multiply-adds, then the Taylor series of sin, cos (to x^9, x^10) and exp (to
 * x^7) in Horner form.  The HIP kernels evaluate the same operations in the
 * same order, so both produce identical bits.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stddef.h>

#include "mc_oracle.h"

#include "oracle_fastmath.h"

void orc_fast_sincosf(float x, float *sn, float *cs) { ofm_sincosf(x, sn, cs); }
float orc_fast_expf(float x) { return ofm_expf(x); }
float orc_fast_logf(float x) { return ofm_logf(x); }

void orc_fast_logf_n(const float *x, int64_t n, float *y) {
    for (int64_t k = 0; k < n; k++) y[k] = orc_fast_logf(x[k]);
}

/* the normal quantile of the Gaussian step for 21-bit k[i] */
void orc_normal_q_n(const uint32_t *k, int64_t cnt, float *y) {
    for (int64_t i = 0; i < cnt; i++) y[i] = ofm_normal_q(k[i]);
}

void orc_fast_sincosf_n(const float *x, int64_t n, float *sn, float *cs) {
    for (int64_t k = 0; k < n; k++) orc_fast_sincosf(x[k], sn + k, cs + k);
}

void orc_fast_expf_n(const float *x, int64_t n, float *y) {
    for (int64_t k = 0; k < n; k++) y[k] = orc_fast_expf(x[k]);
}

/* exp(min(x, 0)) with x clamped to [-87, 0] (normal results only): the fast
 * path accepts iff v < accept_weight(-beta dE), which is "dE <= 0 or
 * v < exp(-beta dE)" for v in [0, 1) */
float orc_accept_weight(float x) {
    x = fminf(fmaxf(x, -87.0f), 0.0f);
    const float k = rintf(x * LOG2E);
    float r = fmaf(-k, LN2_HI, x);
    r = fmaf(-k, LN2_LO, r);
    float p = fmaf(r, E7, E6);
    p = fmaf(p, r, E5);
    p = fmaf(p, r, E4);
    p = fmaf(p, r, E3);
    p = fmaf(p, r, E2);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    union { uint32_t u; float f; } two_k;
    two_k.u = (uint32_t)((int)k + 127) << 23;
    return p * two_k.f;
}

void orc_xyfast_init_cache(const float *th, int64_t n, float *c, float *s) {
    for (int64_t k = 0; k < n; k++) orc_fast_sincosf(th[k], s + k, c + k);
}

#define FSITE(L, r, i, j) ((size_t)(r) * (size_t)(L) * (size_t)(L) + (size_t)(i) * (size_t)(L) + (size_t)(j))

void orc_xyfast_color(float *th, float *c, float *s, uint64_t *rng, const float *betas, int R, int L,
                      int color, float J, float delta) {
    const float one = 1.0f, two = 2.0f;
    const float pi = (float)M_PI, two_pi = (float)(2.0 * M_PI);
    const int h = L / 2;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < R; r++) {
        for (int row = 0; row < h; row++) {
            for (int col = 0; col < h; col++) {
                const int i = 2 * row + color / 2, j = 2 * col + color % 2;
                const size_t k = FSITE(L, r, i, j);
                uint64_t *st = rng + 2 * k;
                const float theta = th[k];
                const float u = orc_uniform_f32(st);
                float p = theta + delta * (two * u - one);
                if (p >= pi) p -= two_pi;
                else if (p < -pi) p += two_pi;
                float sp, cp;
                orc_fast_sincosf(p, &sp, &cp);
                const size_t ku = FSITE(L, r, (i + L - 1) % L, j), kd = FSITE(L, r, (i + 1) % L, j);
                const size_t kl = FSITE(L, r, i, (j + L - 1) % L), kr = FSITE(L, r, i, (j + 1) % L);
                const float C = ((c[ku] + c[kd]) + c[kl]) + c[kr];
                const float S = ((s[ku] + s[kd]) + s[kl]) + s[kr];
                float t = (cp - c[k]) * C;
                t = fmaf(sp - s[k], S, t);
                const float dE = -J * t;
                const float v = orc_uniform_f32(st);
                if (v < orc_accept_weight(-betas[r] * dE)) {
                    th[k] = p;
                    c[k] = cp;
                    s[k] = sp;
                }
            }
        }
    }
}

/* Gaussian step of the fast path from ONE 64-bit draw x (the batched
 * engine's gauss_draw): bits 63..43 -> k, z = +-Q((2k + 1) 2^-23) with Q the
 * normal quantile of ofm_normal_q (inverse-CDF sampling); bit 42 -> the sign;
 * bits 21..1 -> acceptance uniform v = k3 2^-21.  orc_xyfast_gauss_color then
 * proposes p = theta + sigma z wrapped by whole turns (multiplying by
 * 1/(2 pi), branch free), with the fast dE and acceptance.  sig holds each
 * walker's sigma; acc counts accepted moves per walker (may be NULL). */
void orc_fast_gauss_draw(uint64_t x, float *z, float *v) {
    const uint32_t hi = (uint32_t)(x >> 32), lo = (uint32_t)x;
    const float q = ofm_normal_q(hi >> 11);
    *z = (hi & 0x400u) ? -q : q;
    *v = (float)((lo >> 1) & 0x1fffffu) * 0x1p-21f;
}

static float fast_wrap(float p) {
    const float pi = (float)M_PI, two_pi = (float)(2.0 * M_PI);
    p -= two_pi * floorf((p + pi) * 0x1.45f306p-3f);
    p = p >= pi ? p - two_pi : p;
    return p < -pi ? p + two_pi : p;
}

void orc_xyfast_gauss_color(float *th, float *c, float *s, uint64_t *rng, const float *betas, const float *sig,
                            int R, int L, int color, float J, int64_t *acc) {
    const int h = L / 2;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < R; r++) {
        int64_t n_acc = 0;
        for (int row = 0; row < h; row++) {
            for (int col = 0; col < h; col++) {
                const int i = 2 * row + color / 2, j = 2 * col + color % 2;
                const size_t k = FSITE(L, r, i, j);
                uint64_t *st = rng + 2 * k;
                const float theta = th[k];
                float z, v;
                orc_fast_gauss_draw(orc_rng_next(st), &z, &v);
                const float p = fast_wrap(theta + sig[r] * z);
                float sp, cp;
                orc_fast_sincosf(p, &sp, &cp);
                const size_t ku = FSITE(L, r, (i + L - 1) % L, j), kd = FSITE(L, r, (i + 1) % L, j);
                const size_t kl = FSITE(L, r, i, (j + L - 1) % L), kr = FSITE(L, r, i, (j + 1) % L);
                const float C = ((c[ku] + c[kd]) + c[kl]) + c[kr];
                const float S = ((s[ku] + s[kd]) + s[kl]) + s[kr];
                float t = (cp - c[k]) * C;
                t = fmaf(sp - s[k], S, t);
                const float dE = -J * t;
                if (v < orc_accept_weight(-betas[r] * dE)) {
                    th[k] = p;
                    c[k] = cp;
                    s[k] = sp;
                    n_acc++;
                }
            }
        }
        if (acc) acc[r] += n_acc;
    }
}

void orc_xyfast_sweeps(float *th, float *c, float *s, uint64_t *rng, const float *betas, int R, int L,
                       float J, float delta, int64_t nsweeps) {
    for (int64_t n = 0; n < nsweeps; n++)
        for (int col = 0; col < 4; col++) orc_xyfast_color(th, c, s, rng, betas, R, L, col, J, delta);
}

/* energy, mx, my, twist_cos_x, twist_sin_x from the cache; per-site terms in
 * float32 exactly as the kernel forms them.  The energy is the batched
 * engine's integer sum: each site's (cos bond x + cos bond y) rounded to the
 * nearest multiple of 2^-40 (ties to even), added exactly in int64, then
 * -J * 2^-40 * sum in float64.  The other sums are float64. */
void orc_xyfast_observables(const float *c, const float *s, int R, int L, float J, double *out) {
    for (int r = 0; r < R; r++) {
        int64_t ef = 0;
        double mx = 0, my = 0, cx = 0, qx = 0;
        for (int i = 0; i < L; i++)
            for (int j = 0; j < L; j++) {
                const size_t k = FSITE(L, r, i, j);
                const size_t kd = FSITE(L, r, (i + 1) % L, j), kr = FSITE(L, r, i, (j + 1) % L);
                const float bxc = fmaf(s[k], s[kd], c[k] * c[kd]);
                const float byc = fmaf(s[k], s[kr], c[k] * c[kr]);
                const float bxs = fmaf(s[k], c[kd], -(c[k] * s[kd]));
                ef += (int64_t)llrint((double)(bxc + byc) * 0x1p40);
                mx += c[k];
                my += s[k];
                cx += bxc;
                qx += bxs;
            }
        out[0 * R + r] = -(double)J * ((double)ef * 0x1p-40);
        out[1 * R + r] = mx;
        out[2 * R + r] = my;
        out[3 * R + r] = cx;
        out[4 * R + r] = qx;
    }
}
