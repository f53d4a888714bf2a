"""Float64 histogram reweighting (utils/reweighting.py), pinned by exact
enumeration of the L = 4 Ising model.

* single histogram: exact Boltzmann weights at beta_0, reweighted, give the
  exact expectations at beta' (algebraic identity, 1e-12);
* a one-chain multiple-histogram estimate is the single-histogram one;
* Metropolis chains of the oracle at four temperatures, combined by the
  multiple-histogram equations, give <E>, <|M|> and log Z ratios at
  temperatures between them within jackknife errors of the exact values;
* a planted mistake (the reweighting factor's sign flipped) fails the check.
"""

import numpy as np
import pytest

import oracle
from utils.reweighting import MultiHistogram, reweight_run, single_histogram


def _enumerate_l4():
    n = 16
    configs = ((np.arange(2**n)[:, None] >> np.arange(n)[None, :]) & 1).astype(np.int8) * 2 - 1
    s = configs.reshape(-1, 4, 4)
    E = -(s * np.roll(s, -1, 1) + s * np.roll(s, -1, 2)).sum(axis=(1, 2)).astype(float)
    M = s.sum(axis=(1, 2)).astype(float)
    return E, M


def _exact(E, M, beta):
    lw = -beta * (E - E.min())
    w = np.exp(lw)
    Z = w.sum()
    return (w * E).sum() / Z, (w * np.abs(M)).sum() / Z, np.log(Z) - beta * E.min()


def test_single_histogram_is_exact_on_exact_weights():
    E, M = _enumerate_l4()
    b0 = 0.4
    w0 = np.exp(-b0 * (E - E.min()))
    for b in (0.25, 0.4407, 0.6):
        out = single_histogram(E, {"absM": np.abs(M)}, b0, [b], weights=w0)
        eE, eM, _ = _exact(E, M, b)
        np.testing.assert_allclose(out["energy"][0], eE, rtol=1e-12)
        np.testing.assert_allclose(out["absM"][0], eM, rtol=1e-12)


def test_one_chain_multi_histogram_equals_single_histogram():
    g = np.random.default_rng(3)
    E = g.normal(-20, 5, 4000)
    O = g.normal(size=4000)
    mh = MultiHistogram(E[:, None], [0.5], {"o": O[:, None]})
    for b in (0.45, 0.55):
        a = mh.expectations([b])
        s = single_histogram(E, {"o": O}, 0.5, [b])
        np.testing.assert_allclose(a["energy"], s["energy"], rtol=1e-12)
        np.testing.assert_allclose(a["o"], s["o"], rtol=1e-11, atol=1e-14)


def _oracle_chains(temps, sweeps=24000, seed=99):
    L, R = 4, len(temps)
    betas = (1.0 / np.asarray(temps)).astype(np.float32)
    spins = np.ones((R, L, L), np.int8)
    rng = oracle.rng_init(R * L * L, seed)
    oracle.ising_sweeps(spins, rng, betas, 1.0, 500)
    E, M = [], []
    for _ in range(sweeps):
        oracle.ising_sweeps(spins, rng, betas, 1.0, 1)
        m = oracle.ising_measure(spins, 1.0)
        E.append(m[0])
        M.append(np.abs(m[1]))
    return np.asarray(E), np.asarray(M), betas.astype(np.float64)


@pytest.fixture(scope="module")
def chains():
    return _oracle_chains([2.0, 2.3, 2.6, 3.0])


def test_multi_histogram_reproduces_exact_l4_between_ladder_points(chains):
    E, M, betas = chains
    Ee, Me = _enumerate_l4()
    mh = MultiHistogram(E, betas, {"absM": M})
    targets = 1.0 / np.array([2.15, 2.45, 2.8])
    vals, errs = mh.jackknife(targets, 1200)
    for i, b in enumerate(targets):
        eE, eM, logz = _exact(Ee, Me, b)
        assert abs(vals["energy"][i] - eE) < 5 * errs["energy"][i], (b, vals["energy"][i], eE, errs["energy"][i])
        assert abs(vals["absM"][i] - eM) < 5 * errs["absM"][i], (b, vals["absM"][i], eM, errs["absM"][i])
    # free-energy differences from the coldest ladder temperature
    logz0 = _exact(Ee, Me, betas[0])[2]
    for i, b in enumerate(targets):
        exact = _exact(Ee, Me, b)[2] - logz0
        assert abs(vals["log_z"][i] - exact) < 5 * errs["log_z"][i], (b, vals["log_z"][i], exact, errs["log_z"][i])
    np.testing.assert_allclose(mh.log_partition_ratio(targets[1]), vals["log_z"][1], rtol=1e-12)


def test_reweight_run_on_a_run_dictionary(chains):
    E, M, betas = chains
    run = {"samples": {"energy": E, "absM": M},
           "metadata": {"L": 4, "temperatures": 1.0 / betas}}
    out = reweight_run(run, [1 / 2.45], observables=["absM"], block_size=1200)
    eE, _, _ = _exact(*_enumerate_l4(), 1 / 2.45)
    assert abs(out["energy"][0] - eE) < 5 * out["energy_err"][0]
    assert out["specific_heat"][0] > 0 and "absM_err" in out


def test_planted_sign_error_is_detected(chains):
    """Reweighting with the factor exp(+(beta - beta0) E) is far outside the errors."""
    E, M, betas = chains
    Ee, Me = _enumerate_l4()
    b = 1 / 2.45
    k = 1  # the chain at T = 2.3
    good = single_histogram(E[:, k], {}, betas[k], [b])["energy"][0]
    # exp(+(b - b0) E) instead of exp(-(b - b0) E) is the correct rule aimed at 2 b0 - b
    bad = single_histogram(E[:, k], {}, betas[k], [2 * betas[k] - b])["energy"][0]
    eE = _exact(Ee, Me, b)[0]
    spread = np.std(E[:, k]) / np.sqrt(len(E) / 50)
    assert abs(good - eE) < 5 * spread
    assert abs(bad - eE) > 10 * abs(good - eE) and abs(bad - eE) > 5 * spread
