"""Pins of the oracle's float32 fast path (the batched XY engine's arithmetic).

* fast_sincosf / fast_expf against float64 libm: error bounds in float32 ulps
  over dense grids, exact odd/even symmetry, exact special values, monotone exp.
  A wrong polynomial coefficient, reduction constant or quadrant fails these.
* The fast XY update evaluates dE = -J[(cos p - cos t) C + (sin p - sin t) S]
  from cached (cos, sin); mathematically it is the reference update
  (models/xy/update.py:6-54), so a site with frozen neighbours must sample the
  von Mises conditional, and the cache must stay equal to fast_sincosf(theta).
* The fast observables equal the reference XY measurement
  (models/xy/measure.py:17-134) up to float32 rounding of the bond terms.
"""

import numpy as np
import pytest
from scipy import special

import oracle


def _ulp_error(got, ref):
    ref32 = ref.astype(np.float32)
    ulp = np.maximum(np.spacing(np.abs(ref32)).astype(np.float64), float(np.spacing(np.float32(2.0**-24))))
    return np.abs(got.astype(np.float64) - ref) / ulp


def test_fast_sincosf_error_bound_and_symmetry():
    grid = np.linspace(-2 * np.pi, 2 * np.pi, 1_000_001).astype(np.float32)
    rand = np.random.default_rng(3).uniform(-7.0, 7.0, 1_000_000).astype(np.float32)
    x = np.concatenate([grid, rand])
    sn, cs = oracle.fast_sincosf(x)
    xd = x.astype(np.float64)
    assert _ulp_error(sn, np.sin(xd)).max() <= 1.5
    assert _ulp_error(cs, np.cos(xd)).max() <= 1.5
    sn_neg, cs_neg = oracle.fast_sincosf(-x)
    assert np.array_equal(sn_neg, -sn)
    assert np.array_equal(cs_neg, cs)
    s0, c0 = oracle.fast_sincosf(np.array([0.0], np.float32))
    assert s0[0] == 0.0 and c0[0] == 1.0
    norm = sn.astype(np.float64)**2 + cs.astype(np.float64)**2
    assert np.abs(norm - 1).max() < 4e-7


def test_fast_expf_error_bound_monotone_and_special_values():
    x = np.linspace(-103.9, 0.0, 2_000_001).astype(np.float32)
    y = oracle.fast_expf(x)
    ref = np.exp(x.astype(np.float64))
    normal = ref > np.finfo(np.float32).tiny
    assert _ulp_error(y[normal], ref[normal]).max() <= 1.0
    assert np.abs(y[~normal] - ref[~normal]).max() < 2 * np.finfo(np.float32).smallest_subnormal
    assert np.all(np.diff(y) >= 0)
    special_in = np.array([0.0, -104.0, -1000.0, -np.inf, np.nan], np.float32)
    out = oracle.fast_expf(special_in)
    assert out[0] == 1.0
    assert np.all(out[1:4] == 0.0)
    assert np.isnan(out[4])


def test_fast_cache_follows_angles_through_sweeps():
    L, R = 8, 3
    theta = np.random.default_rng(5).uniform(-np.pi, np.pi, (R, L, L)).astype(np.float32)
    c, s = oracle.xyfast_cache(theta)
    rng = oracle.rng_init(R * L * L, 9)
    betas = np.array([0.3, 1.0, 3.0], np.float32)
    oracle.xyfast_sweeps(theta, c, s, rng, betas, 1.0, 1.5, 20)
    s_ref, c_ref = oracle.fast_sincosf(theta)
    assert np.array_equal(c, c_ref) and np.array_equal(s, s_ref)
    assert np.all((theta >= -np.pi) & (theta < np.float32(np.pi)))


def test_fast_observables_equal_reference_measurement():
    L, R, J = 16, 4, 1.3
    theta = np.random.default_rng(7).uniform(-np.pi, np.pi, (R, L, L)).astype(np.float32)
    c, s = oracle.xyfast_cache(theta)
    fast = oracle.xyfast_observables(c, s, J)
    ref = oracle.xy_measure(theta, J)
    assert np.allclose(fast, ref, rtol=0, atol=L * L * 2e-6), np.abs(fast - ref).max()
    # ordered state: every bond at -J, twist sums L^2 and 0
    zero = np.zeros((1, L, L), np.float32)
    c0, s0 = oracle.xyfast_cache(zero)
    e, mx, my, cx, qx = oracle.xyfast_observables(c0, s0, J)[:, 0]
    assert (e, mx, my, cx, qx) == (-2 * float(np.float32(J)) * L * L, L * L, 0.0, L * L, 0.0)


def test_fast_energy_is_an_exact_integer_sum():
    """The batched engines' energy: per-site (cos bond x + cos bond y) rounded to
    multiples of 2^-40 and added as integers, so any summation order -- here the
    lattice translated, which reorders every sum -- gives the same bits."""
    L, J = 32, 1.0
    theta = np.random.default_rng(11).uniform(-np.pi, np.pi, (1, L, L)).astype(np.float32)
    c, s = oracle.xyfast_cache(theta)
    e0 = oracle.xyfast_observables(c, s, J)[0, 0]
    for shift in [(3, 0), (0, 5), (7, 11)]:
        rolled = np.ascontiguousarray(np.roll(theta, shift, axis=(1, 2)))
        cr, sr = oracle.xyfast_cache(rolled)
        assert oracle.xyfast_observables(cr, sr, J)[0, 0] == e0
    # quantisation: within 2^-41 per site of the float64 sum of the float32 terms
    cd, sd = c.astype(np.float64), s.astype(np.float64)
    down = lambda a: np.roll(a, -1, axis=1)
    right = lambda a: np.roll(a, -1, axis=2)
    # fmaf(s, s_n, c * c_n) in float32, as the kernels form the bond terms
    bx = (sd * down(sd) + (c * down(c)).astype(np.float64)).astype(np.float32)
    by = (sd * right(sd) + (c * right(c)).astype(np.float64)).astype(np.float32)
    approx = -J * np.sum((bx + by).astype(np.float64))
    assert abs(e0 - approx) <= L * L * 2.0 ** -41 + 1e-9
    assert abs(e0 - (-J * np.sum(cd * down(cd) + sd * down(sd) + cd * right(cd) + sd * right(sd)))) < L * L * 1e-6


def _binned(x, nbins):
    b = np.asarray(x)[: len(x) // nbins * nbins].reshape(nbins, -1).mean(axis=1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nbins)


def test_fast_update_single_site_conditional_is_von_mises():
    L, R = 4, 2
    theta0 = np.random.default_rng(11).uniform(-np.pi, np.pi, (R, L, L)).astype(np.float32)
    betas = np.array([0.4, 1.5], np.float32)
    states = oracle.rng_init(R * L * L, 8)
    theta = theta0.copy()
    c, s = oracle.xyfast_cache(theta)
    sites = [(i, j) for i in range(0, L, 2) for j in range(0, L, 2)]
    samples = []
    for _ in range(30000):
        oracle.xyfast_color(theta, c, s, states, betas, 0, 1.0, 1.5)
        samples.append([[theta[r, i, j] for (i, j) in sites] for r in range(R)])
    samples = np.asarray(samples, dtype=float)
    for r in range(R):
        for k, (i, j) in enumerate(sites):
            nb = [theta0[r, (i - 1) % L, j], theta0[r, (i + 1) % L, j], theta0[r, i, (j - 1) % L],
                  theta0[r, i, (j + 1) % L]]
            S = np.sum(np.exp(1j * np.asarray(nb, float)))
            kappa, phi = float(betas[r]) * abs(S), np.angle(S)
            x = samples[:, r, k] - phi
            for stat, expect in ((np.cos(x), special.iv(1, kappa) / special.iv(0, kappa)),
                                 (np.sin(x), 0.0),
                                 (np.cos(2 * x), special.iv(2, kappa) / special.iv(0, kappa))):
                mean, err = _binned(stat, 30)
                assert abs(mean - expect) < 5 * err + 1e-3, (r, k, mean, expect, err)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_fast_update_hot_and_cold_limits(sign):
    """beta = 0 accepts every proposal; a huge beta rejects every uphill one."""
    L, R = 8, 2
    theta = np.zeros((R, L, L), np.float32) if sign > 0 else \
        np.random.default_rng(1).uniform(-np.pi, np.pi, (R, L, L)).astype(np.float32)
    c, s = oracle.xyfast_cache(theta)
    rng = oracle.rng_init(R * L * L, 4)
    probe = rng.copy()
    before = theta.copy()
    betas = np.array([0.0, 1e30], np.float32)
    oracle.xyfast_color(theta, c, s, rng, betas, 0, 1.0, 1.5)
    u = oracle.rng_draw_f32(probe, 0, 1)  # first draw of site (0, 0) of replica 0
    p = np.float32(before[0, 0, 0] + np.float32(1.5) * (np.float32(2) * u[0] - np.float32(1)))
    p = p - np.float32(2 * np.pi) if p >= np.float32(np.pi) else (p + np.float32(2 * np.pi) if p < -np.float32(np.pi) else p)
    assert theta[0, 0, 0] == p
    if sign > 0:  # ordered start: every proposal is uphill at beta = 1e30
        assert np.array_equal(theta[1], before[1])


def test_accept_weight_is_clamped_exponential():
    """v < accept_weight(-beta dE) is the Metropolis rule "dE <= 0 or v < exp(-beta dE)"."""
    x = np.linspace(-87.0, 0.0, 200_001).astype(np.float32)
    np.testing.assert_array_equal(oracle.accept_weight(x), oracle.fast_expf(x))
    assert np.all(oracle.accept_weight(np.array([0.0, 1e-30, 3.0, 1e30, np.inf], np.float32)) == 1.0)
    low = oracle.accept_weight(np.array([-87.0, -90.0, -1e4, -np.inf], np.float32))
    assert np.all(low == low[0]) and 0 < low[0] < 2e-38 and low[0] >= np.finfo(np.float32).tiny


def test_fast_logf_error_bound_on_uniform_draws():
    """fast_logf (the float32 log of the fast paths, called on uniforms in (0, 1])
    is within 1.5 ulp of float64 log, exact at 1 and at powers of two."""
    rand = np.random.default_rng(5).uniform(0.0, 1.0, 1_000_000).astype(np.float32)
    grid = (np.arange(1, 2**20 + 1, dtype=np.float64) * 2.0**-20).astype(np.float32)
    x = np.concatenate([rand[rand > 0], grid, np.float32(2.0) ** -np.arange(1, 127, dtype=np.float32)])
    y = oracle.fast_logf(x)
    assert _ulp_error(y, np.log(x.astype(np.float64))).max() <= 1.5
    assert oracle.fast_logf(np.array([1.0], np.float32))[0] == 0.0
    k = np.arange(-126, 1)
    np.testing.assert_array_equal(oracle.fast_logf((2.0 ** k).astype(np.float32)),
                                  (k * np.log(2.0)).astype(np.float32))


def test_normal_quantile_table_every_input():
    """normal_q(k) = Q((2k + 1) 2^-23) with Q = -Phi^-1, the Gaussian step's
    inverse-CDF table, over all 2^21 inputs: within 2.01 ulp of max(Q, 1) of
    scipy's float64 ndtri, positive, and with mean square 1 (the points are the
    midpoints of 2^22 equal-probability bins of (0, 1/2), so the mean of Q^2
    is the variance of the discretised normal)."""
    k = np.arange(2**21, dtype=np.uint32)
    y = oracle.normal_q(k)
    q = -special.ndtri((2.0 * k.astype(np.float64) + 1.0) * 2.0**-23)
    scale = np.spacing(np.maximum(q, 1.0).astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(y - q) / scale) <= 2.01
    assert np.all(y > 0) and 5.16 < y[0] < 5.17  # Q(2^-23)
    assert abs(np.mean(y.astype(np.float64)**2) - 1.0) < 1e-4
    assert abs(np.mean(y.astype(np.float64)**4) - 3.0) < 1e-3
