"""Host-side utilities: run files, temperature ladders, estimators and the
autocorrelation time used to size jackknife blocks.  Pinned by closed forms
(AR(1) autocorrelation, linear estimators, exact line crossings, a
Weber-Minnhagen curve generated from known parameters) and by invariants
(round trips, endpoints, the focus fraction of the ladder)."""

import json

import numpy as np
import pytest

from main.precision_study import block_size_for, tau_int
from utils.analysis import (blocked_jackknife, fit_weber_minnhagen, jackknife_from_blocks, linear_crossings,
                            xy_analysis)
from utils.data_io import load_runs, save_runs
from utils.temperature_ladder import make_temperature_ladder


def _ar1(phi, n, seed):
    rng = np.random.default_rng(seed)
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - phi * phi)
    for i in range(1, n):
        x[i] = phi * x[i - 1] + e[i]
    return x


@pytest.mark.parametrize("phi", [0.0, 0.6, 0.95])
def test_tau_int_matches_ar1_closed_form(phi):
    exact = (1 + phi) / (2 * (1 - phi))
    n = 200_000
    # Madras-Sokal: var(tau) ~ 2 (2W + 1) / n tau^2 with window W ~ 6 tau
    sigma = exact * np.sqrt(2 * (2 * 6 * exact + 1) / n)
    assert abs(tau_int(_ar1(phi, n, 4)) - exact) < 4 * sigma + 0.02


def test_block_size_keeps_ten_blocks_and_covers_tau():
    assert block_size_for({"a": 3.0, "b": 12.0}, 10_000) == 48
    assert block_size_for({"a": 3.0}, 10_000) == 20
    assert block_size_for({"a": 500.0}, 1_000) == 100


def test_jackknife_of_a_mean_is_the_standard_error_of_block_means():
    x = np.random.default_rng(1).normal(size=(1000, 3))
    value, err = blocked_jackknife({"x": x}, lambda m: m["x"], 50)
    blocks = x.reshape(20, 50, 3).mean(axis=1)
    np.testing.assert_allclose(value, blocks.mean(0), rtol=1e-12)
    np.testing.assert_allclose(err, blocks.std(0, ddof=1) / np.sqrt(20), rtol=1e-10)
    with pytest.raises(ValueError):
        jackknife_from_blocks({"x": blocks[:1]}, lambda m: m["x"])


def test_xy_estimators_on_a_frozen_ordered_lattice():
    L, n, T = 8, 400, np.array([0.5, 1.0])
    samples = {"energy": np.full((n, 2), -2.0 * L * L), "mx": np.full((n, 2), float(L * L)),
               "my": np.zeros((n, 2)), "twist_cos_x": np.full((n, 2), float(L * L)),
               "twist_sin_x": np.zeros((n, 2))}
    run = {"samples": samples, "metadata": {"temperatures": T, "model_parameters": {"J": 1.0}}}
    r = xy_analysis(run, L, block_size=40)
    np.testing.assert_allclose(r["energy"], -2.0)
    np.testing.assert_allclose(r["binder"], 0.5)      # 1 - 1/2 for a fixed |m|
    np.testing.assert_allclose(r["helicity"], 1.0)    # J <cx> / L^2, no twist fluctuations
    np.testing.assert_allclose(r["energy_err"], 0.0, atol=1e-14)


def test_run_files_round_trip(tmp_path):
    runs = {16: {"samples": {"energy": np.arange(6.0).reshape(3, 2), "mx": np.ones((3, 2))},
                 "metadata": {"L": np.int64(16), "temperatures": np.array([1.0, 2.0]),
                              "model_parameters": {"J": np.float32(1.5)}, "sample_sweeps": np.array([10, 20, 30])}},
            32: {"samples": {"energy": np.zeros((2, 2))}, "metadata": {"L": 32, "temperatures": [1.0, 2.0]}}}
    path = tmp_path / "sub" / "run.npz"
    save_runs(path, runs)
    back = load_runs(path)
    assert sorted(back) == [16, 32]
    np.testing.assert_array_equal(back[16]["samples"]["energy"], runs[16]["samples"]["energy"])
    assert back[16]["metadata"]["model_parameters"]["J"] == 1.5
    assert back[16]["metadata"]["sample_sweeps"] == [10, 20, 30]
    with np.load(path, allow_pickle=False) as data:
        manifest = json.loads(data["manifest"].item())
    assert manifest["version"] == 1 and manifest["runs"]["32"]["sample_names"] == ["energy"]


def test_uniform_beta_ladder_and_endpoints():
    T = make_temperature_ladder(0.8, 1.2, 9)
    np.testing.assert_allclose(np.diff(1 / T[::-1]), np.full(8, (1 / 0.8 - 1 / 1.2) / 8), rtol=1e-12)
    assert T[0] == pytest.approx(0.8) and T[-1] == pytest.approx(1.2)
    assert np.all(np.diff(T) > 0)


def test_focused_ladder_puts_the_requested_fraction_near_tc():
    T = make_temperature_ladder(0.78, 1.02, 100, Tc=0.893, tc_window=0.08, tc_fraction=0.6)
    assert np.all(np.diff(T) > 0)
    assert T[0] == pytest.approx(0.78) and T[-1] == pytest.approx(1.02)
    inside = np.mean((T > 0.893 - 0.04) & (T < 0.893 + 0.04))
    assert abs(inside - 0.6) < 0.03
    with pytest.raises(ValueError):
        make_temperature_ladder(0.78, 1.02, 10, Tc=1.0, tc_window=0.08)
    with pytest.raises(ValueError):
        make_temperature_ladder(1.0, 0.5, 10)


def test_linear_crossings_are_exact_for_lines():
    T = np.linspace(0.0, 1.0, 11)
    assert linear_crossings(T, 2 * T - 0.7) == [pytest.approx(0.35)]
    assert linear_crossings(T, T, 1 - T) == [pytest.approx(0.5)]
    assert linear_crossings(T, np.ones(11)) == []
    assert np.isnan(linear_crossings(T, np.full(11, np.nan))[0])


def test_weber_minnhagen_fit_recovers_generated_parameters():
    T = np.linspace(0.85, 0.95, 41)
    sizes = np.array([16, 32, 64, 128], float)
    Tc, C = 0.893, 1.2
    # helicity curves through the finite-size jump at Tc: Upsilon(Tc, L) = 2 Tc/pi (1 + 1/(2 ln L + C))
    jump = 2 * Tc / np.pi * (1 + 1 / (2 * np.log(sizes) + C))
    rho = jump[:, None] - 1.5 * (T[None, :] - Tc)
    fit = fit_weber_minnhagen(T, sizes, rho, np.full_like(rho, 1e-3))
    assert fit["temperature"] == pytest.approx(Tc, abs=1e-4)
    assert fit["C"] == pytest.approx(C, abs=1e-2)
    assert fit["chi2"] < 1e-6


def test_crossing_jackknife_recovers_the_crossing_of_noisy_lines():
    from main.fss_study import _crossing_jackknife
    T = np.linspace(0.0, 1.0, 11)
    rng = np.random.default_rng(3)
    B = 40
    a = {"y": 2 * T[None, :] - 0.7 + 0.01 * rng.normal(size=(B, 11))}
    b = {"y": 0.3 + 0.0 * T[None, :] + 0.01 * rng.normal(size=(B, 11))}
    c, err = _crossing_jackknife({"a": a, "b": b}, lambda m: (m["a"]["y"], m["b"]["y"]), T)
    assert abs(c - 0.5) < 5 * err and 0.0005 < err < 0.01   # exact crossing 0.5; error ~ 0.01 / 2 / sqrt(40)
    exact = {"y": np.repeat(2 * T[None, :] - 0.7, B, axis=0)}
    c, err = _crossing_jackknife({"a": exact}, lambda m: (m["a"]["y"], 0.3), T)
    assert c == pytest.approx(0.5) and err == pytest.approx(0.0, abs=1e-12)
