/*
 * Precision-generic part of the CPU oracle (test infrastructure only).
 * Included twice by mc_oracle.c: REAL=float with SUF=_f32, REAL=double with SUF=_f64.
 * Every function follows the reference file:line cited above it.
 */

#define FN(name) CAT(name, SUF)

typedef struct { REAL re, im; } FN(cpx);
#define CPX FN(cpx)

static inline CPX FN(cx)(REAL re, REAL im) { CPX z; z.re = re; z.im = im; return z; }
static inline CPX FN(cadd)(CPX a, CPX b) { return FN(cx)(a.re + b.re, a.im + b.im); }
static inline CPX FN(csub)(CPX a, CPX b) { return FN(cx)(a.re - b.re, a.im - b.im); }
static inline CPX FN(cmul)(CPX a, CPX b) {
    return FN(cx)(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline CPX FN(cconj)(CPX a) { return FN(cx)(a.re, -a.im); }
static inline CPX FN(cscale)(REAL s, CPX a) { return FN(cx)(s * a.re, s * a.im); }
static inline REAL FN(norm2)(CPX z) { return z.re * z.re + z.im * z.im; }

/* site pointer helpers */
#define SITE(L, r, i, j) ((size_t)(r) * (size_t)(L) * (size_t)(L) + (size_t)(i) * (size_t)(L) + (size_t)(j))

/* ------------------------------------------------------------------ XY */

/* models/xy/update.py:6-54 */
void FN(orc_xy_color)(REAL *th, uint64_t *rng, const REAL *betas, int R, int L,
                      int color, REAL J, REAL delta) {
    const REAL one = (REAL)1, two = (REAL)2, zero = (REAL)0;
    const REAL pi = (REAL)M_PI, two_pi = (REAL)(2.0 * M_PI);
    const int h = L / 2;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < R; r++) {
        for (int row = 0; row < h; row++) {
            for (int col = 0; col < h; col++) {
                const int i = 2 * row + color / 2, j = 2 * col + color % 2;
                const size_t k = SITE(L, r, i, j);
                uint64_t *s = rng + 2 * k;
                const REAL theta = th[k];
                const REAL u = FN(draw)(s);
                REAL p = theta + delta * (two * u - one);
                if (p >= pi) p -= two_pi;
                else if (p < -pi) p += two_pi;
                const REAL up = th[SITE(L, r, (i + L - 1) % L, j)];
                const REAL down = th[SITE(L, r, (i + 1) % L, j)];
                const REAL left = th[SITE(L, r, i, (j + L - 1) % L)];
                const REAL right = th[SITE(L, r, i, (j + 1) % L)];
                const REAL e_old = -J * (((FN(rcos)(theta - up) + FN(rcos)(theta - down))
                                          + FN(rcos)(theta - left)) + FN(rcos)(theta - right));
                const REAL e_new = -J * (((FN(rcos)(p - up) + FN(rcos)(p - down))
                                          + FN(rcos)(p - left)) + FN(rcos)(p - right));
                const REAL dE = e_new - e_old;
                const REAL v = FN(draw)(s);
                if (dE <= zero || v < FN(rexp)(-betas[r] * dE)) th[k] = p;
            }
        }
    }
}

/* Gaussian-proposal XY colour update (no reference counterpart; the
 * reference draws a uniform window, models/xy/update.py:24-31): u1, u2 ->
 * z = sqrt(-2 log u1) cos(2 pi u2) (u1 = 0 -> 2^-53), p = theta + sigma z
 * wrapped to [-pi, pi) by whole turns, then the reference dE and acceptance
 * with a third draw.  sigma comes from the ladder temperature whose beta is
 * nearest to the walker's; accepted moves are counted per temperature. */
static inline REAL FN(wrap_angle)(REAL p, REAL pi, REAL two_pi) {
    if (p >= pi || p < -pi) {
        p -= two_pi * FN(rfloor)((p + pi) / two_pi);
        if (p >= pi) p -= two_pi;
        else if (p < -pi) p += two_pi;
    }
    return p;
}

static inline int FN(nearest_temperature)(const REAL *table, int T, REAL beta) {
    int best = 0;
    REAL bd = FN(rfabs)(table[0] - beta);
    for (int t = 1; t < T; t++) {
        const REAL d = FN(rfabs)(table[t] - beta);
        if (d < bd) { bd = d; best = t; }
    }
    return best;
}

void FN(orc_xy_gauss_color)(REAL *th, uint64_t *rng, const REAL *betas, int R, int L, int color, REAL J,
                            const REAL *table_betas, const REAL *sigmas, int T, int64_t *accepted) {
    const REAL two = (REAL)2, zero = (REAL)0;
    const REAL pi = (REAL)M_PI, two_pi = (REAL)(2.0 * M_PI);
    const int h = L / 2;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < R; r++) {
        const int t = FN(nearest_temperature)(table_betas, T, betas[r]);
        const REAL sigma = sigmas[t];
        int64_t acc = 0;
        for (int row = 0; row < h; row++) {
            for (int col = 0; col < h; col++) {
                const int i = 2 * row + color / 2, j = 2 * col + color % 2;
                const size_t k = SITE(L, r, i, j);
                uint64_t *s = rng + 2 * k;
                const REAL theta = th[k];
                REAL u1 = FN(draw)(s);
                const REAL u2 = FN(draw)(s);
                if (u1 <= zero) u1 = (REAL)0x1p-53;
                const REAL z = FN(rsqrt)(-two * FN(rlog)(u1)) * FN(rcos)(two_pi * u2);
                const REAL p = FN(wrap_angle)(theta + sigma * z, pi, two_pi);
                const REAL up = th[SITE(L, r, (i + L - 1) % L, j)];
                const REAL down = th[SITE(L, r, (i + 1) % L, j)];
                const REAL left = th[SITE(L, r, i, (j + L - 1) % L)];
                const REAL right = th[SITE(L, r, i, (j + 1) % L)];
                const REAL e_old = -J * (((FN(rcos)(theta - up) + FN(rcos)(theta - down))
                                          + FN(rcos)(theta - left)) + FN(rcos)(theta - right));
                const REAL e_new = -J * (((FN(rcos)(p - up) + FN(rcos)(p - down))
                                          + FN(rcos)(p - left)) + FN(rcos)(p - right));
                const REAL dE = e_new - e_old;
                const REAL v = FN(draw)(s);
                if (dE <= zero || v < FN(rexp)(-betas[r] * dE)) { th[k] = p; acc++; }
            }
        }
        if (accepted) {
#pragma omp atomic
            accepted[t] += acc;
        }
    }
}

/* core/runtime.py:87-99: colours 0..3 in order, per sweep */
void FN(orc_xy_sweeps)(REAL *th, uint64_t *rng, const REAL *betas, int R, int L,
                       REAL J, REAL delta, int64_t nsweeps) {
    for (int64_t s = 0; s < nsweeps; s++)
        for (int c = 0; c < 4; c++) FN(orc_xy_color)(th, rng, betas, R, L, c, J, delta);
}

/* models/xy/measure.py:17-68 (energy, mx, my, twist_cos_x, twist_sin_x) */
void FN(orc_xy_measure)(const REAL *th, int R, int L, double J, const int32_t *tind,
                        const double *cached, double *out) {
    for (int r = 0; r < R; r++) {
        double e = 0, mx = 0, my = 0, cxs = 0, qx = 0;
        for (int i = 0; i < L; i++) {
            for (int j = 0; j < L; j++) {
                const double theta = (double)th[SITE(L, r, i, j)];
                const double tx = (double)th[SITE(L, r, (i + 1) % L, j)];
                const double dx = theta - tx;
                const double cos_x = cos(dx);
                if (!cached) {
                    const double ty = (double)th[SITE(L, r, i, (j + 1) % L)];
                    e += -J * (cos_x + cos(theta - ty));
                }
                mx += cos(theta);
                my += sin(theta);
                cxs += cos_x;
                qx += sin(dx);
            }
        }
        const int d = tind ? tind[r] : r;
        out[0 * R + d] = cached ? cached[r] : e;
        out[1 * R + d] = mx;
        out[2 * R + d] = my;
        out[3 * R + d] = cxs;
        out[4 * R + d] = qx;
    }
}

/* models/xy/measure.py:71-96 */
void FN(orc_xy_energy)(const REAL *th, int R, int L, double J, double *out) {
    for (int r = 0; r < R; r++) {
        double e = 0;
        for (int i = 0; i < L; i++)
            for (int j = 0; j < L; j++) {
                const double theta = (double)th[SITE(L, r, i, j)];
                const double tx = (double)th[SITE(L, r, (i + 1) % L, j)];
                const double ty = (double)th[SITE(L, r, i, (j + 1) % L)];
                e += -J * (cos(theta - tx) + cos(theta - ty));
            }
        out[r] = e;
    }
}

/* ------------------------------------------------------------ Kawamura */

/* models/kawamura/update.py:7-18; neighbours in order up, down, left, right */
REAL FN(orc_kawamura_local)(const REAL *st, int L, int r, int i, int j, REAL theta,
                            REAL tau, REAL dx, REAL dy, REAL dz, REAL J) {
    const int ni[4] = {(i + L - 1) % L, (i + 1) % L, i, i};
    const int nj[4] = {j, j, (j + L - 1) % L, (j + 1) % L};
    REAL energy = (REAL)0;
    for (int k = 0; k < 4; k++) {
        const REAL *n = st + 5 * SITE(L, r, ni[k], nj[k]);
        const REAL dot = (dx * n[2] + dy * n[3]) + dz * n[4];
        energy -= ((J * ((REAL)1 + tau * n[1])) * dot) * FN(rcos)(theta - n[0]);
    }
    return energy;
}

/* models/kawamura/update.py:21-65: theta move, tau flip, sphere move.  Pinned by the exact
 * single-site conditional (tests/test_oracle_kawamura_conditional.py). */
void FN(orc_kawamura_color)(REAL *st, uint64_t *rng, const REAL *betas, int R, int L,
                            int color, REAL J, REAL delta) {
    const REAL one = (REAL)1, two = (REAL)2, zero = (REAL)0;
    const REAL pi = (REAL)M_PI, two_pi = (REAL)(2.0 * M_PI);
    const int h = L / 2;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < R; r++) {
        for (int row = 0; row < h; row++) {
            for (int col = 0; col < h; col++) {
                const int i = 2 * row + color / 2, j = 2 * col + color % 2;
                const size_t k = SITE(L, r, i, j);
                uint64_t *s = rng + 2 * k;
                REAL *site = st + 5 * k;
                const REAL beta = betas[r];
                REAL theta = site[0], tau = site[1], dx = site[2], dy = site[3], dz = site[4];
                REAL energy = FN(orc_kawamura_local)(st, L, r, i, j, theta, tau, dx, dy, dz, J);

                REAL pt = theta + delta * (two * FN(draw)(s) - one);
                if (pt >= pi) pt -= two_pi;
                else if (pt < -pi) pt += two_pi;
                REAL pe = FN(orc_kawamura_local)(st, L, r, i, j, pt, tau, dx, dy, dz, J);
                REAL dE = pe - energy;
                REAL u = FN(draw)(s);
                if (dE <= zero || u < FN(rexp)(-beta * dE)) { theta = pt; energy = pe; }

                const REAL ptau = -tau;
                pe = FN(orc_kawamura_local)(st, L, r, i, j, theta, ptau, dx, dy, dz, J);
                dE = pe - energy;
                u = FN(draw)(s);
                if (dE <= zero || u < FN(rexp)(-beta * dE)) { tau = ptau; energy = pe; }

                const REAL z = two * FN(draw)(s) - one;
                const REAL phi = two_pi * FN(draw)(s);
                const REAL radius = FN(rsqrt)(one - z * z);
                const REAL x = radius * FN(rcos)(phi), y = radius * FN(rsin)(phi);
                pe = FN(orc_kawamura_local)(st, L, r, i, j, theta, tau, x, y, z, J);
                dE = pe - energy;
                u = FN(draw)(s);
                if (dE <= zero || u < FN(rexp)(-beta * dE)) { dx = x; dy = y; dz = z; }

                site[0] = theta; site[1] = tau; site[2] = dx; site[3] = dy; site[4] = dz;
            }
        }
    }
}

void FN(orc_kawamura_sweeps)(REAL *st, uint64_t *rng, const REAL *betas, int R, int L,
                             REAL J, REAL delta, int64_t nsweeps) {
    for (int64_t s = 0; s < nsweeps; s++)
        for (int c = 0; c < 4; c++) FN(orc_kawamura_color)(st, rng, betas, R, L, c, J, delta);
}

/* models/kawamura/measure.py:14-70
 * (energy, chirality, px, py, spin2, twist_cos_x, twist_sin_x) */
void FN(orc_kawamura_measure)(const REAL *st, int R, int L, double J, const int32_t *tind,
                              const double *cached, double *out) {
    for (int r = 0; r < R; r++) {
        double e = 0, chir = 0, px = 0, py = 0, cxs = 0, qx = 0;
        double xx = 0, yy = 0, zz = 0, xy = 0, xz = 0, yz = 0;
        for (int i = 0; i < L; i++) {
            for (int j = 0; j < L; j++) {
                const REAL *s = st + 5 * SITE(L, r, i, j);
                const double theta = s[0], tau = s[1], dx = s[2], dy = s[3], dz = s[4];
                chir += tau;
                px += cos(2 * theta);
                py += sin(2 * theta);
                xx += dx * dx - 1.0 / 3;
                yy += dy * dy - 1.0 / 3;
                zz += dz * dz - 1.0 / 3;
                xy += dx * dy;
                xz += dx * dz;
                yz += dy * dz;
                const int dirs = cached ? 1 : 2;
                for (int d = 0; d < dirs; d++) {
                    const int ni = d == 0 ? (i + 1) % L : i;
                    const int nj = d == 0 ? j : (j + 1) % L;
                    const REAL *n = st + 5 * SITE(L, r, ni, nj);
                    const double dot = (dx * (double)n[2] + dy * (double)n[3]) + dz * (double)n[4];
                    const double weight = (1 + tau * (double)n[1]) * dot;
                    const double angle = theta - (double)n[0];
                    const double cosine = weight * cos(angle);
                    if (!cached) e -= J * cosine;
                    if (d == 0) {
                        cxs += cosine;
                        qx += weight * sin(angle);
                    }
                }
            }
        }
        const int d = tind ? tind[r] : r;
        out[0 * R + d] = cached ? cached[r] : e;
        out[1 * R + d] = chir;
        out[2 * R + d] = px;
        out[3 * R + d] = py;
        out[4 * R + d] = (xx * xx + yy * yy + zz * zz + 2 * (xy * xy + xz * xz + yz * yz)) / 3;
        out[5 * R + d] = cxs;
        out[6 * R + d] = qx;
    }
}

/* models/kawamura/measure.py:71-99 */
void FN(orc_kawamura_energy)(const REAL *st, int R, int L, double J, double *out) {
    for (int r = 0; r < R; r++) {
        double energy = 0;
        for (int i = 0; i < L; i++)
            for (int j = 0; j < L; j++) {
                const REAL *s = st + 5 * SITE(L, r, i, j);
                const double theta = s[0], tau = s[1];
                for (int d = 0; d < 2; d++) {
                    const int ni = d == 0 ? (i + 1) % L : i;
                    const int nj = d == 0 ? j : (j + 1) % L;
                    const REAL *n = st + 5 * SITE(L, r, ni, nj);
                    double dot = 0;
                    for (int c = 2; c < 5; c++) dot += (double)s[c] * (double)n[c];
                    energy -= J * (1 + tau * (double)n[1]) * dot * cos(theta - (double)n[0]);
                }
            }
        out[r] = energy;
    }
}

/* ----------------------------------------------------------------- He3 */

static inline CPX FN(ld)(const REAL *p) { return FN(cx)(p[0], p[1]); }
/* component (mu, a) of site k; state layout (R, L, L, 3, 2) complex, re/im interleaved */
#define HE3(st, L, r, i, j, mu, a) ((st) + 2 * (6 * SITE(L, r, i, j) + 2 * (mu) + (a)))

/* models/he3/update.py:14-27 */
static REAL FN(dipole_invariant)(const CPX *psi, REAL n) {
    const CPX a = psi[0], b = psi[1], c = psi[2], d = psi[3];
    return (((FN(norm2)(FN(cadd)(a, d)) + FN(norm2)(a)) + FN(norm2)(d))
            + (REAL)2 * FN(cmul)(b, FN(cconj)(c)).re)
           - (REAL)(2.0 / 3.0) * n;
}

/* models/he3/update.py:31-61; psi[mu*2 + a] */
static REAL FN(potential)(const CPX *psi, const REAL *b, REAL g_d) {
    REAL nx = FN(norm2)(psi[0]), ny = FN(norm2)(psi[1]);
    CPX xx = FN(cmul)(psi[0], psi[0]), yy = FN(cmul)(psi[1], psi[1]);
    CPX xy = FN(cmul)(psi[0], psi[1]);
    CPX overlap = FN(cmul)(FN(cconj)(psi[0]), psi[1]);
    for (int mu = 1; mu < 3; mu++) {
        const CPX x = psi[2 * mu], y = psi[2 * mu + 1];
        nx += FN(norm2)(x);
        ny += FN(norm2)(y);
        xx = FN(cadd)(xx, FN(cmul)(x, x));
        yy = FN(cadd)(yy, FN(cmul)(y, y));
        xy = FN(cadd)(xy, FN(cmul)(x, y));
        overlap = FN(cadd)(overlap, FN(cmul)(FN(cconj)(x), y));
    }
    const REAL n = nx + ny, two = (REAL)2;
    const REAL i1 = FN(norm2)(FN(cadd)(xx, yy));
    const REAL i2 = n * n;
    const REAL i3 = (nx * nx + ny * ny) + two * FN(cmul)(overlap, overlap).re;
    const REAL i4 = (nx * nx + ny * ny) + two * FN(norm2)(overlap);
    const REAL i5 = (FN(norm2)(xx) + FN(norm2)(yy)) + two * FN(norm2)(xy);
    const REAL dip = FN(dipole_invariant)(psi, n);
    return (((((-two * n + b[0] * i1) + b[1] * i2) + b[2] * i3) + b[3] * i4) + b[4] * i5)
           + g_d * dip;
}

REAL FN(orc_he3_potential)(const REAL *psi_flat, const REAL *coef, REAL g_d) {
    CPX psi[6];
    for (int k = 0; k < 6; k++) psi[k] = FN(ld)(psi_flat + 2 * k);
    return FN(potential)(psi, coef, g_d);
}

/* models/he3/update.py:64-70 */
static REAL FN(gradient)(CPX dx_x, CPX dx_y, CPX dy_x, CPX dy_y) {
    return (((REAL)3 * (FN(norm2)(dx_x) + FN(norm2)(dy_y)) + FN(norm2)(dx_y)) + FN(norm2)(dy_x))
           + (REAL)2 * (FN(cmul)(FN(cconj)(dx_x), dy_y).re + FN(cmul)(FN(cconj)(dx_y), dy_x).re);
}

/* models/he3/update.py:73-91 */
REAL FN(orc_he3_local_energy)(const REAL *st, int L, int r, int i, int j, const REAL *psi_flat,
                              const REAL *coef, REAL K, REAL g_d) {
    const int ip = (i + 1) % L, im = (i + L - 1) % L, jp = (j + 1) % L, jm = (j + L - 1) % L;
    CPX psi[6];
    for (int k = 0; k < 6; k++) psi[k] = FN(ld)(psi_flat + 2 * k);
    REAL g = (REAL)0;
    for (int mu = 0; mu < 3; mu++) {
        const CPX x = psi[2 * mu], y = psi[2 * mu + 1];
        const CPX lx = FN(ld)(HE3(st, L, r, im, j, mu, 0)), ly = FN(ld)(HE3(st, L, r, im, j, mu, 1));
        const CPX dx = FN(ld)(HE3(st, L, r, i, jm, mu, 0)), dy = FN(ld)(HE3(st, L, r, i, jm, mu, 1));
        g += FN(gradient)(FN(csub)(FN(ld)(HE3(st, L, r, ip, j, mu, 0)), x),
                          FN(csub)(FN(ld)(HE3(st, L, r, ip, j, mu, 1)), y),
                          FN(csub)(FN(ld)(HE3(st, L, r, i, jp, mu, 0)), x),
                          FN(csub)(FN(ld)(HE3(st, L, r, i, jp, mu, 1)), y));
        g += FN(gradient)(FN(csub)(x, lx), FN(csub)(y, ly),
                          FN(csub)(FN(ld)(HE3(st, L, r, im, jp, mu, 0)), lx),
                          FN(csub)(FN(ld)(HE3(st, L, r, im, jp, mu, 1)), ly));
        g += FN(gradient)(FN(csub)(FN(ld)(HE3(st, L, r, ip, jm, mu, 0)), dx),
                          FN(csub)(FN(ld)(HE3(st, L, r, ip, jm, mu, 1)), dy),
                          FN(csub)(x, dx), FN(csub)(y, dy));
    }
    return FN(potential)(psi, coef, g_d) + K * g;
}

/* models/he3/update.py:93-173 (exact change of the three affected gradient cells) */
static REAL FN(gradient_difference)(const REAL *st, int L, int r, int i, int j, const CPX *prop) {
    const int ip = (i + 1) % L, im = (i + L - 1) % L, jp = (j + 1) % L, jm = (j + L - 1) % L;
    const REAL three = (REAL)3, two = (REAL)2, eight = (REAL)8, four = (REAL)4;
    REAL delta_g = (REAL)0;
    for (int mu = 0; mu < 3; mu++) {
        const CPX x = FN(ld)(HE3(st, L, r, i, j, mu, 0)), y = FN(ld)(HE3(st, L, r, i, j, mu, 1));
        const CPX ex = FN(csub)(prop[2 * mu], x), ey = FN(csub)(prop[2 * mu + 1], y);
        const CPX rx = FN(ld)(HE3(st, L, r, ip, j, mu, 0)), ry = FN(ld)(HE3(st, L, r, ip, j, mu, 1));
        const CPX ux = FN(ld)(HE3(st, L, r, i, jp, mu, 0)), uy = FN(ld)(HE3(st, L, r, i, jp, mu, 1));
        const CPX lx = FN(ld)(HE3(st, L, r, im, j, mu, 0)), ly = FN(ld)(HE3(st, L, r, im, j, mu, 1));
        const CPX dx = FN(ld)(HE3(st, L, r, i, jm, mu, 0)), dy = FN(ld)(HE3(st, L, r, i, jm, mu, 1));
        const CPX lux = FN(ld)(HE3(st, L, r, im, jp, mu, 0)), luy = FN(ld)(HE3(st, L, r, im, jp, mu, 1));
        const CPX drx = FN(ld)(HE3(st, L, r, ip, jm, mu, 0)), dry = FN(ld)(HE3(st, L, r, ip, jm, mu, 1));
        const CPX a = FN(csub)(rx, x), b = FN(csub)(ry, y), c = FN(csub)(ux, x), d = FN(csub)(uy, y);
        const CPX al = FN(csub)(x, lx), bl = FN(csub)(y, ly), cl = FN(csub)(lux, lx), dl = FN(csub)(luy, ly);
        const CPX ad = FN(csub)(drx, dx), bd = FN(csub)(dry, dy), cd = FN(csub)(x, dx), dd = FN(csub)(y, dy);
        CPX t = FN(cadd)(FN(cadd)(FN(cadd)(FN(cscale)(three, a), b), c), d);
        const CPX fx = FN(cadd)(FN(cadd)(FN(cx)(-t.re, -t.im), FN(cadd)(FN(cscale)(three, al), dl)),
                                FN(cadd)(cd, bd));
        t = FN(cadd)(FN(cadd)(FN(cadd)(a, b), c), FN(cscale)(three, d));
        const CPX fy = FN(cadd)(FN(cadd)(FN(cx)(-t.re, -t.im), FN(cadd)(bl, cl)),
                                FN(cadd)(FN(cscale)(three, dd), ad));
        const REAL linear = two * (FN(cmul)(FN(cconj)(ex), fx).re + FN(cmul)(FN(cconj)(ey), fy).re);
        const REAL quadratic = eight * (FN(norm2)(ex) + FN(norm2)(ey))
                               + four * FN(cmul)(FN(cconj)(ex), ey).re;
        delta_g += linear + quadratic;
    }
    return delta_g;
}

REAL FN(orc_he3_gradient_difference)(const REAL *st, int L, int r, int i, int j,
                                     const REAL *proposed) {
    CPX prop[6];
    for (int k = 0; k < 6; k++) prop[k] = FN(ld)(proposed + 2 * k);
    return FN(gradient_difference)(st, L, r, i, j, prop);
}

/* models/he3/update.py:176-227 (Cartesian box proposal, 12 draws + 1) */
void FN(orc_he3_color)(REAL *st, uint64_t *rng, const REAL *betas, int R, int L, int color,
                       const REAL *coef, REAL K, REAL delta, REAL g_d) {
    const int h = L / 2;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < R; r++) {
        for (int row = 0; row < h; row++) {
            for (int col = 0; col < h; col++) {
                const int i = 2 * row + color / 2, j = 2 * col + color % 2;
                const size_t k = SITE(L, r, i, j);
                uint64_t *s = rng + 2 * k;
                CPX psi[6];
                for (int m = 0; m < 6; m++) psi[m] = FN(ld)(HE3(st, L, r, i, j, m / 2, m % 2));
                const REAL old_potential = FN(potential)(psi, coef, g_d);
                for (int m = 0; m < 6; m++) {
                    const REAL re = delta * ((REAL)2 * FN(draw)(s) - (REAL)1);
                    const REAL im = delta * ((REAL)2 * FN(draw)(s) - (REAL)1);
                    psi[m] = FN(cx)(psi[m].re + re, psi[m].im + im);
                }
                const REAL dpot = FN(potential)(psi, coef, g_d) - old_potential;
                const REAL dgrad = FN(gradient_difference)(st, L, r, i, j, psi);
                const REAL diff = dpot + K * dgrad;
                const REAL u = FN(draw)(s);
                if (diff <= (REAL)0 || u < FN(rexp)(-betas[r] * diff)) {
                    for (int m = 0; m < 6; m++) {
                        REAL *dst = HE3(st, L, r, i, j, m / 2, m % 2);
                        dst[0] = psi[m].re;
                        dst[1] = psi[m].im;
                    }
                }
            }
        }
    }
}

void FN(orc_he3_sweeps)(REAL *st, uint64_t *rng, const REAL *betas, int R, int L,
                        const REAL *coef, REAL K, REAL delta, REAL g_d, int64_t nsweeps) {
    for (int64_t s = 0; s < nsweeps; s++)
        for (int c = 0; c < 4; c++) FN(orc_he3_color)(st, rng, betas, R, L, c, coef, K, delta, g_d);
}

/* models/he3/gaussian_fast.py:26-40 */
static void FN(neighbor_field)(const REAL *st, int L, int r, int i, int j, CPX *field) {
    const int ip = (i + 1) % L, im = (i + L - 1) % L, jp = (j + 1) % L, jm = (j + L - 1) % L;
    const REAL three = (REAL)3;
    for (int mu = 0; mu < 3; mu++) {
        const CPX rx = FN(ld)(HE3(st, L, r, ip, j, mu, 0)), ry = FN(ld)(HE3(st, L, r, ip, j, mu, 1));
        const CPX ux = FN(ld)(HE3(st, L, r, i, jp, mu, 0)), uy = FN(ld)(HE3(st, L, r, i, jp, mu, 1));
        const CPX lx = FN(ld)(HE3(st, L, r, im, j, mu, 0)), ly = FN(ld)(HE3(st, L, r, im, j, mu, 1));
        const CPX dx = FN(ld)(HE3(st, L, r, i, jm, mu, 0)), dy = FN(ld)(HE3(st, L, r, i, jm, mu, 1));
        const CPX lux = FN(ld)(HE3(st, L, r, im, jp, mu, 0)), luy = FN(ld)(HE3(st, L, r, im, jp, mu, 1));
        const CPX drx = FN(ld)(HE3(st, L, r, ip, jm, mu, 0)), dry = FN(ld)(HE3(st, L, r, ip, jm, mu, 1));
        /* (-three*(rx+lx)-ux-uy-ry-ly-dx-dy+luy+dry), left to right */
        CPX f = FN(cscale)(-three, FN(cadd)(rx, lx));
        f = FN(csub)(f, ux); f = FN(csub)(f, uy); f = FN(csub)(f, ry); f = FN(csub)(f, ly);
        f = FN(csub)(f, dx); f = FN(csub)(f, dy); f = FN(cadd)(f, luy); f = FN(cadd)(f, dry);
        field[2 * mu] = f;
        /* (-rx-ry-ux-lx-ly-dx-three*(uy+dy)+lux+drx) */
        CPX g = FN(cx)(-rx.re, -rx.im);
        g = FN(csub)(g, ry); g = FN(csub)(g, ux); g = FN(csub)(g, lx); g = FN(csub)(g, ly);
        g = FN(csub)(g, dx); g = FN(csub)(g, FN(cscale)(three, FN(cadd)(uy, dy)));
        g = FN(cadd)(g, lux); g = FN(cadd)(g, drx);
        field[2 * mu + 1] = g;
    }
}

/* models/he3/gaussian_fast.py:43-50 (Box-Muller pair) */
static CPX FN(normal_pair)(uint64_t *s) {
    REAL u = FN(draw)(s);
    if (u <= (REAL)0) u = (REAL)1.1102230246251565e-16;
    const REAL angle = (REAL)6.283185307179586 * FN(draw)(s);
    const REAL radius = FN(rsqrt)(-(REAL)2 * FN(rlog)(u));
    return FN(cx)(radius * FN(rcos)(angle), radius * FN(rsin)(angle));
}

/* models/he3/gaussian_fast.py:53-88; overwrites psi, returns log q(rev)/q(fwd) */
static REAL FN(gaussian_proposal)(CPX *psi, const CPX *field, REAL beta, REAL K, REAL rho,
                                  REAL mass, uint64_t *s) {
    const REAL one = (REAL)1, two = (REAL)2;
    const REAL root = (REAL)0.7071067811865475;
    const REAL plus = (REAL)10 * K + mass, minus = (REAL)6 * K + mass;
    const REAL vf = one - rho * rho;
    const REAL sigma_plus = FN(rsqrt)(vf / (two * beta * plus));
    const REAL sigma_minus = FN(rsqrt)(vf / (two * beta * minus));
    const REAL q_plus = vf / ((((REAL)8 * root) * root) * sigma_plus * sigma_plus);
    const REAL q_minus = vf / ((((REAL)8 * root) * root) * sigma_minus * sigma_minus);
    REAL log_ratio = (REAL)0;
    for (int mu = 0; mu < 3; mu++) {
        const CPX x = psi[2 * mu], y = psi[2 * mu + 1];
        const CPX hp = FN(cscale)(root, FN(cadd)(field[2 * mu], field[2 * mu + 1]));
        const CPX hm = FN(cscale)(root, FN(csub)(field[2 * mu], field[2 * mu + 1]));
        /* mp = -K*hp/plus ; mm = -K*hm/minus (complex / real) */
        const CPX kp = FN(cscale)(-K, hp), km = FN(cscale)(-K, hm);
        const CPX mp = FN(cx)(kp.re / plus, kp.im / plus);
        const CPX mm = FN(cx)(km.re / minus, km.im / minus);
        const CPX mx = FN(cscale)(root, FN(cadd)(mp, mm));
        const CPX my = FN(cscale)(root, FN(csub)(mp, mm));
        const CPX zp = FN(cscale)(sigma_plus, FN(normal_pair)(s));
        const CPX zm = FN(cscale)(sigma_minus, FN(normal_pair)(s));
        const CPX nx = FN(cadd)(FN(cadd)(mx, FN(cscale)(rho, FN(csub)(x, mx))),
                                FN(cscale)(root, FN(cadd)(zp, zm)));
        const CPX ny = FN(cadd)(FN(cadd)(my, FN(cscale)(rho, FN(csub)(y, my))),
                                FN(cscale)(root, FN(csub)(zp, zm)));
        psi[2 * mu] = nx;
        psi[2 * mu + 1] = ny;
        const CPX old_plus = FN(cadd)(FN(csub)(x, mx), FN(csub)(y, my));
        const CPX old_minus = FN(csub)(FN(csub)(x, mx), FN(csub)(y, my));
        const CPX new_plus = FN(cadd)(FN(csub)(nx, mx), FN(csub)(ny, my));
        const CPX new_minus = FN(csub)(FN(csub)(nx, mx), FN(csub)(ny, my));
        log_ratio += q_plus * (FN(norm2)(new_plus) - FN(norm2)(old_plus));
        log_ratio += q_minus * (FN(norm2)(new_minus) - FN(norm2)(old_minus));
    }
    return log_ratio;
}

/* models/he3/gaussian_fast.py:91-119 */
void FN(orc_he3_gaussian_color)(REAL *st, uint64_t *rng, const REAL *betas, int R, int L,
                                int color, const REAL *coef, REAL K, REAL rho, REAL mass,
                                int64_t *counts, REAL g_d) {
    const int h = L / 2;
    const size_t N = (size_t)R * L * L;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < R; r++) {
        for (int row = 0; row < h; row++) {
            for (int col = 0; col < h; col++) {
                const int i = 2 * row + color / 2, j = 2 * col + color % 2;
                const size_t k = SITE(L, r, i, j);
                uint64_t *s = rng + 2 * k;
                CPX field[6], psi[6];
                FN(neighbor_field)(st, L, r, i, j, field);
                for (int m = 0; m < 6; m++) psi[m] = FN(ld)(HE3(st, L, r, i, j, m / 2, m % 2));
                const REAL old_potential = FN(potential)(psi, coef, g_d);
                const REAL lq = FN(gaussian_proposal)(psi, field, betas[r], K, rho, mass, s);
                const REAL dh = (FN(potential)(psi, coef, g_d) - old_potential)
                                + K * FN(gradient_difference)(st, L, r, i, j, psi);
                const REAL log_ratio = -betas[r] * dh + lq;
                const REAL u = FN(draw)(s);
                counts[0 * N + k] += 1;
                if (!isfinite(log_ratio)) {
                    counts[2 * N + k] += 1;
                } else if (log_ratio >= (REAL)0 || u < FN(rexp)(log_ratio)) {
                    counts[1 * N + k] += 1;
                    for (int m = 0; m < 6; m++) {
                        REAL *dst = HE3(st, L, r, i, j, m / 2, m % 2);
                        dst[0] = psi[m].re;
                        dst[1] = psi[m].im;
                    }
                }
            }
        }
    }
}

/* models/he3/measure.py:22-89 (10 channels) */
void FN(orc_he3_measure)(const REAL *st, int R, int L, const REAL *coef, REAL K, REAL g_d,
                         const int32_t *tind, const double *cached, double *out) {
    for (int r = 0; r < R; r++) {
        double energy = 0, density = 0, chir = 0, pr = 0, pim = 0, ix = 0, kx = 0, iy = 0, ky = 0, sz = 0;
        for (int i = 0; i < L; i++) {
            for (int j = 0; j < L; j++) {
                CPX psi[6];
                for (int m = 0; m < 6; m++) psi[m] = FN(ld)(HE3(st, L, r, i, j, m / 2, m % 2));
                if (!cached) energy += (double)FN(potential)(psi, coef, g_d);
                for (int mu = 0; mu < 3; mu++) {
                    const CPX x = psi[2 * mu], y = psi[2 * mu + 1];
                    const CPX xp_x = FN(ld)(HE3(st, L, r, (i + 1) % L, j, mu, 0));
                    const CPX xp_y = FN(ld)(HE3(st, L, r, (i + 1) % L, j, mu, 1));
                    const CPX yp_x = FN(ld)(HE3(st, L, r, i, (j + 1) % L, mu, 0));
                    const CPX yp_y = FN(ld)(HE3(st, L, r, i, (j + 1) % L, mu, 1));
                    const CPX a = FN(csub)(xp_x, x), b = FN(csub)(xp_y, y);
                    const CPX c = FN(csub)(yp_x, x), d = FN(csub)(yp_y, y);
                    if (!cached) energy += (double)(K * FN(gradient)(a, b, c, d));
                    const REAL row_density = FN(norm2)(x) + FN(norm2)(y);
                    density += (double)row_density;
                    if (mu == 2) sz += (double)row_density;
                    chir += 2.0 * (double)FN(cmul)(FN(cconj)(x), y).im;
                    const CPX phase = FN(csub)(FN(cmul)(x, x), FN(cmul)(y, y));
                    pr += (double)phase.re;
                    pim += (double)phase.im;
                    /* integer-literal products promote to double, as in the reference */
                    const double ar = a.re, ai = -a.im, dr = d.re, di = -d.im;
                    const double t1r = 3.0 * ar + dr, t1i = 3.0 * ai + di;
                    const CPX bc = FN(cadd)(FN(cconj)(b), FN(cconj)(c));
                    const CPX bcx = FN(cmul)(bc, xp_y);
                    const double zxr = (t1r * (double)xp_x.re - t1i * (double)xp_x.im) + (double)bcx.re;
                    const double zxi = (t1r * (double)xp_x.im + t1i * (double)xp_x.re) + (double)bcx.im;
                    const CPX cb = FN(cadd)(FN(cconj)(c), FN(cconj)(b));
                    const CPX cby = FN(cmul)(cb, yp_x);
                    const double t2r = 3.0 * dr + ar, t2i = 3.0 * di + ai;
                    const double zyr = (double)cby.re + (t2r * (double)yp_y.re - t2i * (double)yp_y.im);
                    const double zyi = (double)cby.im + (t2r * (double)yp_y.im + t2i * (double)yp_y.re);
                    const double twoK = 2.0 * (double)K;
                    ix += twoK * zxi;
                    kx += twoK * ((3.0 * (double)FN(norm2)(xp_x) + (double)FN(norm2)(xp_y)) - zxr);
                    iy += twoK * zyi;
                    ky += twoK * (((double)FN(norm2)(yp_x) + 3.0 * (double)FN(norm2)(yp_y)) - zyr);
                }
            }
        }
        const int d = tind ? tind[r] : r;
        out[0 * R + d] = cached ? cached[r] : energy;
        out[1 * R + d] = density;
        out[2 * R + d] = chir;
        out[3 * R + d] = pr;
        out[4 * R + d] = pim;
        out[5 * R + d] = ix;
        out[6 * R + d] = kx;
        out[7 * R + d] = iy;
        out[8 * R + d] = ky;
        out[9 * R + d] = density == 0.0 ? 0.0 : 0.5 * (3 * sz / density - 1);
    }
}

/* models/he3/measure.py:115-143 */
void FN(orc_he3_energy)(const REAL *st, int R, int L, const REAL *coef, REAL K, REAL g_d,
                        double *out) {
    for (int r = 0; r < R; r++) {
        double energy = 0;
        for (int i = 0; i < L; i++)
            for (int j = 0; j < L; j++) {
                CPX psi[6];
                for (int m = 0; m < 6; m++) psi[m] = FN(ld)(HE3(st, L, r, i, j, m / 2, m % 2));
                energy += (double)FN(potential)(psi, coef, g_d);
                for (int mu = 0; mu < 3; mu++) {
                    const CPX a = FN(csub)(FN(ld)(HE3(st, L, r, (i + 1) % L, j, mu, 0)), psi[2 * mu]);
                    const CPX b = FN(csub)(FN(ld)(HE3(st, L, r, (i + 1) % L, j, mu, 1)), psi[2 * mu + 1]);
                    const CPX c = FN(csub)(FN(ld)(HE3(st, L, r, i, (j + 1) % L, mu, 0)), psi[2 * mu]);
                    const CPX d = FN(csub)(FN(ld)(HE3(st, L, r, i, (j + 1) % L, mu, 1)), psi[2 * mu + 1]);
                    energy += (double)(K * FN(gradient)(a, b, c, d));
                }
            }
        out[r] = energy;
    }
}

/* ------------------------------------------------------ PT exchange */

/* core/parallel_tempering.py:35-91 (pairs of one parity, then round-trip bookkeeping) */
void FN(orc_pt_exchange)(const double *E, REAL *betas, int32_t *w2t, int32_t *t2w,
                         uint64_t *pt_rng, int64_t *counts, int64_t *trips, int R,
                         int parity, int64_t sweep) {
    const int P = R - 1;
    for (int left = parity; left < R - 1; left += 2) {
        const int right = left + 1;
        const int a = t2w[left], b = t2w[right];
        const double ea = E[a], eb = E[b];
        counts[0 * P + left] += 1;
        if (!isfinite(ea) || !isfinite(eb)) {
            counts[2 * P + left] += 1;
            continue;
        }
        const double log_ratio = ((double)betas[a] - (double)betas[b]) * (ea - eb);
        const double u = orc_uniform_f64(pt_rng + 2 * left);
        if (log_ratio >= 0.0 || u == 0.0 || log(u) < log_ratio) {
            t2w[left] = b;
            t2w[right] = a;
            w2t[a] = right;
            w2t[b] = left;
            const REAL tmp = betas[a];
            betas[a] = betas[b];
            betas[b] = tmp;
            counts[1 * P + left] += 1;
        }
    }
    for (int w = 0; w < R; w++) {
        const int t = w2t[w];
        if (t != 0 && t != R - 1) continue;
        const int64_t origin = trips[0 * R + w];
        if (origin < 0) {
            trips[0 * R + w] = t;
            trips[2 * R + w] = sweep;
        } else if (t != origin) {
            trips[1 * R + w] = 1;
        } else if (trips[1 * R + w]) {
            trips[3 * R + w] += 1;
            trips[4 * R + w] += sweep - trips[2 * R + w];
            trips[2 * R + w] = sweep;
            trips[1 * R + w] = 0;
        }
    }
}

#undef CPX
#undef FN
