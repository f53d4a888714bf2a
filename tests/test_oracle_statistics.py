"""Closed-form pins of the oracle's Markov chains (what the mathematics fixes).

* Ising: Kaufman's exact partition function of the periodic L x L lattice
  (Kaufman 1949; Ferdinand & Fisher 1969), itself checked against brute-force
  enumeration at L = 4, then used as the exact energy for the PT run of the
  L = 32, R = 8 ladder 2.0-2.6 (2,000 + 10,000 sweeps, exchanges every sweep).
* XY: with the other colours frozen, a site's angle must sample the von Mises
  conditional exp(kappa cos(theta - phi)); <cos(theta - phi)> = I1/I0(kappa).
* He3: the conditional-Gaussian MH sampler and the Cartesian Metropolis
  sampler target the same distribution, so their mean energies agree.
A dropped term, wrong sign or wrong beta in an acceptance test fails these.
"""

import numpy as np
import pytest
from scipy import special

import oracle


def _brute_force_ising(L, betas):
    n = L * L
    configs = ((np.arange(2**n)[:, None] >> np.arange(n)[None, :]) & 1).astype(np.int8) * 2 - 1
    s = configs.reshape(-1, L, L)
    E = -(s * np.roll(s, -1, 1) + s * np.roll(s, -1, 2)).sum(axis=(1, 2)).astype(float)
    M = s.sum(axis=(1, 2)).astype(float)
    out = []
    for b in betas:
        w = np.exp(-b * (E - E.min()))
        Z = w.sum()
        out.append(((w * E).sum() / Z, (w * np.abs(M)).sum() / Z, (w * M**2).sum() / Z, np.log(Z) - b * E.min()))
    return np.array(out)


def kaufman_log_z(L, K):
    """log Z of the L x L periodic Ising model at coupling K = beta J."""
    r = np.arange(L)
    def gamma(l):
        c = np.cosh(2 * K) / np.tanh(2 * K) - np.cos(np.pi * l / L)
        return np.arccosh(c)
    g_odd = gamma(2 * r + 1)
    g_even = gamma(2 * r).astype(float)
    g_even[0] = 2 * K + np.log(np.tanh(K))  # signed gamma_0
    terms = []
    for g, fn in ((g_odd, "cosh"), (g_odd, "sinh"), (g_even, "cosh"), (g_even, "sinh")):
        x = 0.5 * L * g
        if fn == "cosh":
            logs = np.abs(x) + np.log1p(np.exp(-2 * np.abs(x)))
            sign = 1.0
        else:
            ax = np.abs(x)
            logs = ax + np.log(-np.expm1(-2 * ax))
            sign = np.prod(np.sign(x))
        terms.append((sign, np.sum(logs)))  # product of 2cosh/2sinh: the 2s cancel the 1/2
    top = max(t[1] for t in terms)
    total = sum(s * np.exp(v - top) for s, v in terms)
    return 0.5 * L * L * np.log(2 * np.sinh(2 * K)) + top + np.log(total) - np.log(2)


def kaufman_energy(L, beta, h=1e-6):
    return -(kaufman_log_z(L, beta + h) - kaufman_log_z(L, beta - h)) / (2 * h)


def test_kaufman_formula_matches_brute_force_enumeration():
    betas = [0.2, 0.4406868, 0.7]
    exact = _brute_force_ising(4, betas)
    for b, (E, _, _, logz) in zip(betas, exact):
        np.testing.assert_allclose(kaufman_log_z(4, b), logz, rtol=1e-12)
        np.testing.assert_allclose(kaufman_energy(4, b), E, rtol=1e-7)


def _binned(x, nbins=20):
    x = np.asarray(x)
    n = len(x) // nbins * nbins
    b = x[:n].reshape(nbins, -1, *x.shape[1:]).mean(axis=1)
    return b.mean(axis=0), b.std(axis=0, ddof=1) / np.sqrt(nbins)


def test_ising_l4_metropolis_samples_the_exact_distribution():
    L, temps = 4, np.array([1.5, 2.27, 3.5])
    betas = (1 / temps).astype(np.float32)
    exact = _brute_force_ising(L, betas.astype(float))
    spins = np.ones((3, L, L), np.int8)
    rng = oracle.rng_init(3 * L * L, 2024)
    oracle.ising_sweeps(spins, rng, betas, 1.0, 1000)
    E, absM, M2 = [], [], []
    for _ in range(40000):
        oracle.ising_sweeps(spins, rng, betas, 1.0, 1)
        m = oracle.ising_measure(spins, 1.0)
        E.append(m[0]); absM.append(np.abs(m[1])); M2.append(m[1]**2)
    for values, col in ((E, 0), (absM, 1), (M2, 2)):
        mean, err = _binned(values)
        z = (mean - exact[:, col]) / np.maximum(err, 1e-12)
        assert np.all(np.abs(z) < 5), (col, mean, exact[:, col], err)


@pytest.mark.slow
def test_ising_l32_pt_ladder_energy_matches_kaufman():
    """The ladder of the L = 32 Ising benchmark: T = linspace(2.0, 2.6, 8),
    seeded random start, exchanges every sweep, 2,000 + 10,000 sweeps."""
    from models.ising.model import IsingModel
    L, R = 32, 8
    temps = np.linspace(2.0, 2.6, R)
    betas = (1.0 / temps).astype(np.float32)
    spins = IsingModel(start="random").initialize_state(L, R, 42 + L)
    rng = oracle.rng_init(R * L * L, 42 + L)
    pt_rng = oracle.rng_init(R - 1, 43)
    w2t = np.arange(R, dtype=np.int32)
    t2w = np.arange(R, dtype=np.int32)
    counts = np.zeros((3, R - 1), np.int64)
    trips = np.zeros((5, R), np.int64)
    trips[0] = trips[2] = -1
    trips[0, 0], trips[0, -1] = 0, R - 1
    trips[2, [0, -1]] = 0
    samples = []
    parity = 0
    for sweep in range(1, 12001):
        oracle.ising_sweeps(spins, rng, betas, 1.0, 1)
        e = oracle.ising_energy(spins, 1.0)
        oracle.pt_exchange(e, betas, w2t, t2w, pt_rng, counts, trips, parity, sweep)
        parity = 1 - parity
        if sweep > 2000:
            ordered = np.empty(R)
            ordered[w2t] = e
            samples.append(ordered / L**2)
    mean, err = _binned(samples, 10)
    exact = np.array([kaufman_energy(L, 1.0 / t) for t in temps]) / L**2
    z = (mean - exact) / err
    assert np.all(np.abs(z) < 5), (mean, exact, err)
    assert np.all(counts[1] > 0)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_xy_single_site_conditional_is_von_mises(dtype):
    L, R = 4, 2
    rng = np.random.default_rng(11)
    theta0 = rng.uniform(-np.pi, np.pi, (R, L, L)).astype(dtype)
    betas = np.array([0.4, 1.5], dtype)
    states = oracle.rng_init(R * L * L, 8)
    theta = theta0.copy()
    sites = [(i, j) for i in range(0, L, 2) for j in range(0, L, 2)]
    samples = []
    for _ in range(30000):
        oracle.xy_color(theta, states, betas, 0, 1.0, 1.5)
        samples.append([[theta[r, i, j] for (i, j) in sites] for r in range(R)])
    samples = np.asarray(samples, dtype=float)
    for r in range(R):
        for s, (i, j) in enumerate(sites):
            nb = [theta0[r, (i - 1) % L, j], theta0[r, (i + 1) % L, j], theta0[r, i, (j - 1) % L],
                  theta0[r, i, (j + 1) % L]]
            S = np.sum(np.exp(1j * np.asarray(nb, float)))
            kappa, phi = float(betas[r]) * abs(S), np.angle(S)
            x = samples[:, r, s] - phi
            for stat, expect in ((np.cos(x), special.iv(1, kappa) / special.iv(0, kappa)),
                                 (np.sin(x), 0.0),
                                 (np.cos(2 * x), special.iv(2, kappa) / special.iv(0, kappa))):
                mean, err = _binned(stat, 30)
                assert abs(mean - expect) < 5 * err + 1e-3, (r, s, mean, expect, err)


def test_he3_gaussian_and_cartesian_samplers_agree():
    from models.he3.model import He3Model
    model = He3Model(pressure=20, g_d=0.001)
    L, R = 4, 2
    betas = (1 / np.array([0.35, 0.6])).astype(np.float32)
    energies = {}
    for sampler in ("cartesian", "gaussian"):
        state = model.initialize_state(L, R, 0)
        rng = oracle.rng_init(R * L * L, 5 if sampler == "cartesian" else 6)
        counts = np.zeros((3, R, L, L), np.int64)
        trace = []
        for sweep in range(24000):
            for color in range(4):
                if sampler == "cartesian":
                    oracle.he3_color(state, rng, betas, color, model.betas, model.K, np.float32(0.3), model.g_d)
                else:
                    oracle.he3_gaussian_color(state, rng, betas, color, model.betas, model.K,
                                              np.float32(-0.9), np.float32(0), counts, model.g_d)
            if sweep >= 4000:
                trace.append(oracle.he3_energy(state, model.betas, model.K, model.g_d))
        energies[sampler] = _binned(trace, 20)
        if sampler == "gaussian":
            assert counts[2].sum() == 0 and counts[1].sum() > 0
    (m1, e1), (m2, e2) = energies["cartesian"], energies["gaussian"]
    z = (m1 - m2) / np.sqrt(e1**2 + e2**2)
    assert np.all(np.abs(z) < 5), (m1, m2, e1, e2)
