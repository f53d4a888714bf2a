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


@pytest.mark.parametrize("kind,L", [("kawamura", 16), ("he3", 12)])
def test_film_models_float32_and_float64_agree_within_errors(gpu, kind, L):
    """Kawamura (chirality, phase and spin order) and He3 (chirality, phase,
    density, helicity): float32 and float64 drop-in chains with on-device
    exchanges agree within combined errors.  The film sizes (L = 192 and 128,
    64 replicas) are main/precision_models.py."""
    from main.precision_models import run
    rep = run((kind,), R=8, thermalization=3000, measurement=12000, sample_interval=10, seed=31,
              sizes={kind: L})
    for name, d in rep["models"][kind]["estimators"].items():
        if d["max_abs_z"] is None:
            continue
        assert d["max_abs_z"] < 4.5, (name, d["z"])
        assert d["mean_z2"] < 3.0, (name, d["z"])
