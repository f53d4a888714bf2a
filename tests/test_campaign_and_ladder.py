"""Campaign output handling and the temperature ladder (host logic, CPU).

* preflight refuses duplicate and existing outputs before any run
  (main/he3.py:90-110, main/kawamura.py:81-93 of the reference) and creates
  nothing;
* save_atomically stages in the destination directory and links the final
  name: the file loads back, an existing file is never overwritten, and no
  staging file is left behind (main/he3.py:192-201, main/kawamura.py:156-165);
* make_temperature_ladder (utils/temperature_ladder.py) against the ladder
  the mathematics of its construction fixes, computed independently: the
  density 1 + A p(beta) with p the tanh window integrated in closed form
  (s log cosh), A from the exact in-window fraction, and quantiles found by
  root finding, for two focused argument sets and the uniform one.
"""

import numpy as np
import pytest
from scipy import optimize

from main._campaign import preflight, save_atomically
from utils.data_io import load_runs
from utils.temperature_ladder import make_temperature_ladder


def _run(L=4):
    return {"samples": {"energy": np.arange(6.0).reshape(3, 2)},
            "metadata": {"L": L, "temperatures": np.array([0.5, 0.7]), "seed": 1,
                         "sample_sweeps": np.array([2, 4, 6])}}


def test_preflight_refuses_duplicates_and_existing_and_creates_nothing(tmp_path):
    a, b = tmp_path / "x" / "L=4.npz", tmp_path / "x" / "L=6.npz"
    assert preflight([a, b]) == [a, b]
    assert not (tmp_path / "x").exists()
    with pytest.raises(ValueError, match="Duplicate simulation output paths"):
        preflight([a, b, a])
    with pytest.raises(ValueError, match=r"\(check sizes\)"):
        preflight([a, a], "Duplicate simulation output paths (check sizes)")
    b.parent.mkdir(parents=True)
    b.write_bytes(b"old")
    with pytest.raises(FileExistsError, match="L=6.npz"):
        preflight([a, b])
    assert not a.exists()


def test_save_atomically_round_trip_and_never_overwrites(tmp_path):
    path = tmp_path / "campaign" / "P=20_gd=0.001" / "L=4.npz"
    save_atomically(path, 4, _run())
    loaded = load_runs(path)
    np.testing.assert_array_equal(loaded[4]["samples"]["energy"], _run()["samples"]["energy"])
    np.testing.assert_array_equal(loaded[4]["metadata"]["sample_sweeps"], [2, 4, 6])
    before = path.read_bytes()
    with pytest.raises(FileExistsError):
        save_atomically(path, 4, _run(4))
    assert path.read_bytes() == before
    assert sorted(p.name for p in path.parent.iterdir()) == ["L=4.npz"]  # no staging left behind


def _analytic_ladder(T_min, T_max, n_T, Tc, tc_window, tc_fraction):
    bmin, bmax = 1 / T_max, 1 / T_min
    b_lo, b_hi = 1 / (Tc + tc_window / 2), 1 / (Tc - tc_window / 2)
    s = (b_hi - b_lo) / 8

    def lncosh(x):
        x = np.abs(x)
        return x + np.log1p(np.exp(-2 * x)) - np.log(2)

    def P(b):  # integral of the tanh window from bmin
        f = lambda x: 0.5 * s * (lncosh((x - b_lo) / s) - lncosh((x - b_hi) / s))
        return f(b) - f(bmin)

    total = P(bmax)
    inside = P(b_hi) - P(b_lo)
    A = (tc_fraction * (bmax - bmin) - (b_hi - b_lo)) / (inside - tc_fraction * total)
    cdf = lambda b: (b - bmin) + A * P(b)
    full = cdf(bmax)
    # the defining property: a fraction tc_fraction of the density lies in the window
    assert abs(((b_hi - b_lo) + A * inside) / full - tc_fraction) < 1e-12
    q = np.linspace(0, 1, n_T)
    betas = [bmin] + [optimize.brentq(lambda b: cdf(b) / full - x, bmin, bmax, xtol=1e-15) for x in q[1:-1]] + [bmax]
    return (1 / np.asarray(betas))[::-1]


@pytest.mark.parametrize("args", [
    (0.15, 0.45, 100, 0.257, 0.06, 0.60),    # the reference's He3 benchmark ladder
    (0.365, 0.415, 64, 0.385, 0.01, 0.35),   # the He3 film ladder of the precision study
])
def test_focused_ladder_matches_its_analytic_construction(args):
    T_min, T_max, n_T, Tc, w, f = args
    got = make_temperature_ladder(T_min, T_max, n_T, Tc=Tc, tc_window=w, tc_fraction=f)
    want = _analytic_ladder(*args)
    assert got.shape == (n_T,)
    assert np.all(np.diff(got) > 0)
    assert got[0] == pytest.approx(T_min, rel=1e-15) and got[-1] == pytest.approx(T_max, rel=1e-15)
    # the construction integrates on a 4097-point grid: O(h^2) from the closed form
    np.testing.assert_allclose(got, want, rtol=2e-6)
    inside = np.sum((got >= Tc - w / 2) & (got <= Tc + w / 2))
    assert abs(inside / n_T - f) <= 2.0 / n_T


def test_uniform_ladder_and_errors():
    T = make_temperature_ladder(0.65, 0.80, 64)
    np.testing.assert_allclose(np.diff(1 / T), np.diff(1 / T)[0], rtol=1e-12)
    assert T[0] == pytest.approx(0.65) and T[-1] == pytest.approx(0.80)
    with pytest.raises(ValueError, match="0 < T_min < T_max"):
        make_temperature_ladder(0.8, 0.6, 4)
    with pytest.raises(ValueError, match="at least two"):
        make_temperature_ladder(0.6, 0.8, 1)
    with pytest.raises(ValueError, match="inside range"):
        make_temperature_ladder(0.6, 0.8, 8, Tc=0.79, tc_window=0.1)
    with pytest.raises(ValueError, match="between 0 and 1"):
        make_temperature_ladder(0.6, 0.8, 8, Tc=0.7, tc_window=0.05, tc_fraction=1.0)
    with pytest.raises(ValueError, match="not achievable"):
        make_temperature_ladder(0.6, 0.8, 8, Tc=0.7, tc_window=0.05, tc_fraction=0.01)
