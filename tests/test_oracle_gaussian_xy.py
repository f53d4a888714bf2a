"""Pins of the oracle's Gaussian-proposal XY updates (plain float32/float64 and
the batched engine's float32 fast restatement).  The reference has no Gaussian
XY proposal, so what pins it is the mathematics:

* at beta = 0 every proposal is accepted, so one colour pass moves each site by
  a wrapped normal N(0, sigma^2): E cos d = exp(-sigma^2/2),
  E cos 2d = exp(-2 sigma^2), E sin d = 0 (Box-Muller and the wrap);
* with frozen neighbours the site samples the von Mises conditional
  (symmetric proposal + Metropolis acceptance);
* sigma follows the walker's temperature through permuted betas, angles stay
  in [-pi, pi) for any sigma, and the counters equal the moves made.
"""

import numpy as np
import pytest
from scipy import special

import oracle


def _wrapped(d):
    return (d + np.pi) % (2 * np.pi) - np.pi


def _binned(x, nbins=30):
    b = np.asarray(x)[: len(x) // nbins * nbins].reshape(nbins, -1).mean(axis=1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nbins)


def _run_plain(theta, states, betas, table, sigmas, color, acc):
    oracle.xy_gauss_color(theta, states, betas, color, 1.0, table, sigmas, acc)


def _run_fast(theta, states, betas, table, sigmas, color, acc, cache):
    # the fast restatement takes each walker's sigma (nearest ladder beta)
    t = np.abs(table[None, :] - betas[:, None]).argmin(axis=1)
    per_walker = np.zeros(len(betas), dtype=np.int64)
    oracle.xyfast_gauss_color(theta, cache[0], cache[1], states, betas, sigmas[t], color, 1.0, per_walker)
    np.add.at(acc, t, per_walker)


@pytest.mark.parametrize("kind", ["f32", "f64", "fast"])
def test_hot_limit_moves_are_wrapped_normal_and_follow_temperature(kind):
    dtype = np.float64 if kind == "f64" else np.float32
    L, R = 32, 3
    sigmas = np.array([0.3, 1.1, 2.6], dtype)
    # walker r sits at temperature perm[r]; distinct, effectively infinite temperatures
    table = np.array([0.0, 1e-30, 2e-30], dtype)
    perm = np.array([2, 0, 1])
    betas = table[perm].copy()
    theta = np.random.default_rng(3).uniform(-np.pi, np.pi, (R, L, L)).astype(dtype)
    states = oracle.rng_init(R * L * L, 17)
    acc = np.zeros(R, dtype=np.int64)
    cache = oracle.xyfast_cache(theta) if kind == "fast" else None
    d_all = [[] for _ in range(R)]
    for _ in range(40):
        before = theta.copy()
        if kind == "fast":
            _run_fast(theta, states, betas, table, sigmas, 0, acc, cache)
        else:
            _run_plain(theta, states, betas, table, sigmas, 0, acc)
        for r in range(R):
            d_all[r].append(_wrapped(theta[r, 0::2, 0::2].astype(float) - before[r, 0::2, 0::2]).ravel())
    assert np.all((theta >= -np.pi) & (theta < dtype(np.pi)))
    assert np.array_equal(acc, np.full(R, 40 * (L // 2) ** 2))  # beta = 0 accepts everything
    for r in range(R):
        d = np.concatenate(d_all[r])
        s = float(sigmas[perm[r]])
        for stat, expect in ((np.cos(d), np.exp(-s * s / 2)), (np.cos(2 * d), np.exp(-2 * s * s)),
                             (np.sin(d), 0.0)):
            mean, err = _binned(stat)
            assert abs(mean - expect) < 5 * err + 2e-3, (kind, r, mean, expect, err)


@pytest.mark.parametrize("kind", ["f32", "f64", "fast"])
def test_frozen_neighbours_give_von_mises_conditional(kind):
    dtype = np.float64 if kind == "f64" else np.float32
    L, R = 4, 2
    theta0 = np.random.default_rng(11).uniform(-np.pi, np.pi, (R, L, L)).astype(dtype)
    table = np.array([0.4, 1.5], dtype)
    betas = table.copy()
    sigmas = np.array([2.0, 0.8], dtype)
    states = oracle.rng_init(R * L * L, 8)
    theta = theta0.copy()
    cache = oracle.xyfast_cache(theta) if kind == "fast" else None
    acc = np.zeros(R, dtype=np.int64)
    sites = [(i, j) for i in range(0, L, 2) for j in range(0, L, 2)]
    samples = []
    for _ in range(20000):
        if kind == "fast":
            _run_fast(theta, states, betas, table, sigmas, 0, acc, cache)
        else:
            _run_plain(theta, states, betas, table, sigmas, 0, acc)
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
                mean, err = _binned(stat)
                assert abs(mean - expect) < 5 * err + 1e-3, (kind, r, k, mean, expect, err)


def test_counters_equal_moves_and_huge_sigma_stays_wrapped():
    L, R = 16, 2
    theta = np.random.default_rng(2).uniform(-np.pi, np.pi, (R, L, L)).astype(np.float32)
    table = np.array([1.0, 2.0], np.float32)
    sigmas = np.array([4 * np.pi, 0.5], np.float32)
    states = oracle.rng_init(R * L * L, 5)
    acc = np.zeros(R, np.int64)
    moved = np.zeros(R, np.int64)
    for color in range(4):
        before = theta.copy()
        oracle.xy_gauss_color(theta, states, table, color, 1.0, table, sigmas, acc)
        moved += (theta != before).reshape(R, -1).sum(1)
    assert np.array_equal(acc, moved)
    assert np.all((theta >= -np.pi) & (theta < np.float32(np.pi)))
    assert 0 < acc[0] < L * L and 0 < acc[1] < L * L


def test_fast_gauss_draw_moments_and_exact_symmetry():
    """One 64-bit draw gives z ~ N(0, 1) (inverse CDF) and v ~ U[0, 1); flipping
    the sign bit (bit 42) negates z exactly and leaves v."""
    n = 100_000
    x = np.random.default_rng(9).integers(0, 2**63, size=n, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    z, v = oracle.fast_gauss_draw(x)
    zd = z.astype(float)
    assert abs(zd.mean()) < 5 / np.sqrt(n)
    assert abs(zd.var() - 1) < 5 * np.sqrt(2 / n)
    assert abs(np.mean(zd**4) - 3) < 5 * np.sqrt(96 / n)
    assert abs(np.mean(np.abs(zd) > 1.959964) - 0.05) < 5 * np.sqrt(0.05 * 0.95 / n)
    assert np.all((v >= 0) & (v < 1)) and abs(v.mean() - 0.5) < 5 * np.sqrt(1 / 12 / n)
    z2, v2 = oracle.fast_gauss_draw(x ^ np.uint64(1 << 42))
    assert np.array_equal(z2, -z) and np.array_equal(v2, v)


def test_fast_gauss_draw_is_the_normal_quantile_of_its_bit_fields():
    x = np.random.default_rng(21).integers(0, 2**63, size=20_000, dtype=np.uint64) * np.uint64(2)
    z, v = oracle.fast_gauss_draw(x)
    xi = x.astype(object)
    k = np.array([(int(a) >> 43) & 0x1FFFFF for a in xi], float)
    sign = np.array([(int(a) >> 42) & 1 for a in xi])
    k3 = np.array([(int(a) >> 1) & 0x1FFFFF for a in xi], float)
    q = -special.ndtri((2 * k + 1) * 2.0**-23)
    ref = np.where(sign, -q, q)
    scale = np.spacing(np.maximum(q, 1.0).astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(z - ref) / scale) <= 2.01
    np.testing.assert_array_equal(v, (k3 * 2.0**-21).astype(np.float32))
