"""Float32 batched sweeps against the float64 drop-in path on the same physics
(Gaussian-MH XY with shared adapted widths, exchanges every 10 sweeps):
E/N, |m|, the Binder cumulant and the helicity modulus agree within their
combined blocked-jackknife errors at every temperature.  The full-size study
(L = 128, R = 32, 50,000 measurement sweeps) is main/precision_study.py."""

import pytest

pytestmark = pytest.mark.gpu


def test_float32_batched_and_float64_plain_agree_within_errors(gpu):
    from main.precision_study import run
    rep = run(L=32, R=8, thermalization=2000, measurement=12000, sample_interval=10, seed=23,
              adapt_sweeps=1000, verbose=False)
    for name, d in rep["estimators"].items():
        assert d["max_abs_z"] < 4.5, (name, d["z"])
        assert d["chi2_per_T"] < 3.0, (name, d["z"])
    assert all(0.4 < a < 0.6 for a in rep["move_acceptance_float32"])
