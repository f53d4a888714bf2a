"""GPU parity at the benchmark shapes, against the oracle on the same seeds.

* XY L = 128, R = 32 (the headline workload): the batched engine with
  Gaussian proposals at the benchmark launch shape (strips of 16 rows,
  512-thread workgroups, 256 workgroups, pair-local exchanges every 10
  sweeps among 32 walkers), 60 sweeps in one launch and again as six
  launches of one exchange each; trajectories, RNG streams, betas, maps,
  swap counts, round trips and acceptance counters bit-identical to the
  oracle's loop (its fast Gaussian update + the reference exchange).
* XY L = 128, R = 32 on the drop-in path in float64 with exchanges every 10
  sweeps: 100 sweeps within 1e-12 per site of the oracle.
* Kawamura L = 192, R = 64 (float32) and the He3 Gaussian sampler at
  128 x 128, R = 64 (complex64, per-site counts) on the drop-in path with
  on-device exchanges: bit-identical to the oracle, with some swaps accepted.
* He3 Gaussian in complex128 over 100 sweeps within 1e-12 per component.
"""

import os

import numpy as np
import pytest

import oracle
from gpu_helpers import device_rng_as_pairs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _oracle_threads():
    oracle.set_threads(min(16, os.cpu_count() or 1))
    yield
    oracle.set_threads(1)


class _PT:
    """The oracle's exchange bookkeeping (core/parallel_tempering.py:44-91)."""

    def __init__(self, R, betas, seed):
        self.betas = betas
        self.w2t = np.arange(R, dtype=np.int32)
        self.t2w = np.arange(R, dtype=np.int32)
        self.counts = np.zeros((3, R - 1), np.int64)
        self.trips = np.zeros((5, R), np.int64)
        self.trips[0] = self.trips[2] = -1
        self.trips[0, 0], self.trips[0, -1] = 0, R - 1
        self.trips[2, [0, -1]] = 0
        self.rng = oracle.rng_init(R - 1, seed)
        self.parity = 0
        self.R = R

    def exchange(self, energies, sweep):
        oracle.pt_exchange(energies, self.betas, self.w2t, self.t2w, self.rng, self.counts, self.trips,
                           self.parity, sweep)
        if self.R > 2:
            self.parity = 1 - self.parity


SIGMAS_C2 = np.geomspace(0.92, 1.28, 32)  # widths of the order the adaptation reaches


def _c2_oracle(nsweeps, seed=170, pt_seed=171, interval=10):
    from models.xy.model import XYModel
    L, R = 128, 32
    temps = np.geomspace(0.7, 1.1, R)
    model = XYModel(start="random", proposal="gaussian", sigma=SIGMAS_C2)
    theta = model.initialize_state(L, R, seed).copy()
    c, s = oracle.xyfast_cache(theta)
    rng = oracle.rng_init(R * L * L, seed)
    pt = _PT(R, (1.0 / temps).astype(np.float32), pt_seed)
    acc_t = np.zeros(R, np.int64)
    sig = SIGMAS_C2.astype(np.float32)
    for sweep in range(1, nsweeps + 1):
        acc_w = np.zeros(R, np.int64)
        for color in range(4):
            oracle.xyfast_gauss_color(theta, c, s, rng, pt.betas, sig[pt.w2t], color, 1.0, acc_w)
        np.add.at(acc_t, pt.w2t, acc_w)
        if sweep % interval == 0:
            pt.exchange(oracle.xyfast_observables(c, s, 1.0)[0], sweep)
    return model, temps, theta, rng, pt, acc_t


@pytest.fixture(scope="module")
def c2_reference():
    return _c2_oracle(60)


@pytest.mark.parametrize("launch_sweeps", [60, 10])
def test_batched_xy_gauss_benchmark_shape_matches_oracle(gpu, c2_reference, launch_sweeps):
    from core.batched import BatchedRun
    model, temps, theta, rng, pt, acc_t = c2_reference
    br = BatchedRun(model, 128, temps, seed=170, exchange_interval=10, pt_seed=171, strip_rows=16, threads=512)
    assert br.R * br.strips == 256
    for _ in range(60 // launch_sweeps):
        br.run(launch_sweeps)
        w2t = br.walker_to_temperature.copy_to_host()
        np.testing.assert_array_equal(np.sort(w2t), np.arange(32))
        np.testing.assert_array_equal(br.temperature_to_walker.copy_to_host()[w2t], np.arange(32))
    assert pt.counts[1].sum() > 0
    np.testing.assert_array_equal(br.snapshot(), theta)
    np.testing.assert_array_equal(device_rng_as_pairs(br.rng_states), rng)
    np.testing.assert_array_equal(br.betas.copy_to_host(), pt.betas)
    np.testing.assert_array_equal(br.walker_to_temperature.copy_to_host(), pt.w2t)
    np.testing.assert_array_equal(br.temperature_to_walker.copy_to_host(), pt.t2w)
    np.testing.assert_array_equal(device_rng_as_pairs(br._pt_rng), pt.rng)
    np.testing.assert_array_equal(br._swap_counts.copy_to_host(), pt.counts)
    np.testing.assert_array_equal(br._trips.copy_to_host(), pt.trips)
    np.testing.assert_array_equal(br._accepted.copy_to_host().astype(np.int64), acc_t)


def test_plain_fp64_xy_benchmark_shape_within_1e12(gpu):
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from models.xy.model import XYModel
    L, R, seed, interval, n = 128, 32, 170, 10, 100
    temps = np.geomspace(0.7, 1.1, R)
    model = XYModel(dtype=np.float64, start="random", proposal="gaussian", sigma=SIGMAS_C2)
    rt = Runtime(model, L, temps, seed=seed)
    ptg = ParallelTempering(rt, exchange_interval=interval, seed=seed + 1)
    for _ in range(n):
        ptg.sweep()
    theta = model.initialize_state(L, R, seed).copy()
    rng = oracle.rng_init(R * L * L, seed)
    table = 1.0 / temps
    pt = _PT(R, table.copy(), seed + 1)
    acc = np.zeros(R, np.int64)
    for sweep in range(1, n + 1):
        for color in range(4):
            oracle.xy_gauss_color(theta, rng, pt.betas, color, 1.0, table, SIGMAS_C2, acc)
        if sweep % interval == 0:
            pt.exchange(oracle.xy_energy(theta, 1.0), sweep)
    assert pt.counts[1].sum() > 0
    np.testing.assert_array_equal(ptg.walker_to_temperature.copy_to_host(), pt.w2t)
    d = np.angle(np.exp(1j * (rt.snapshot() - theta)))
    assert np.abs(d).max() < 1e-12, np.abs(d).max()


def test_kawamura_film_shape_bit_identical_with_exchanges(gpu):
    from core.parallel_tempering import ParallelTempering
    from main.precision_models import runtime, temperatures
    L, R, seed, interval, n = 192, 64, 234, 2, 6
    temps = temperatures("kawamura", R)
    rt = runtime("kawamura", L, temps, False, seed)
    ptg = ParallelTempering(rt, exchange_interval=interval, seed=seed + 1)
    for _ in range(n):
        ptg.sweep()
    model = rt.model
    state = model.initialize_state(L, R, seed).copy()
    rng = oracle.rng_init(R * L * L, seed)
    pt = _PT(R, (1.0 / temps).astype(np.float32), seed + 1)
    for sweep in range(1, n + 1):
        for color in range(4):
            oracle.kawamura_color(state, rng, pt.betas, color, model.J, model.delta)
        if sweep % interval == 0:
            pt.exchange(oracle.kawamura_energy(state, model.J), sweep)
    assert pt.counts[1].sum() > 0
    np.testing.assert_array_equal(rt.snapshot(), state)
    np.testing.assert_array_equal(device_rng_as_pairs(rt.rng_states), rng)
    np.testing.assert_array_equal(ptg.walker_to_temperature.copy_to_host(), pt.w2t)
    np.testing.assert_array_equal(rt.betas.copy_to_host(), pt.betas)
    d = ptg.diagnostics()
    np.testing.assert_array_equal(d["attempts"], pt.counts[0])
    np.testing.assert_array_equal(d["accepted"], pt.counts[1])


def test_he3_gaussian_film_shape_bit_identical_with_counts(gpu):
    from core.parallel_tempering import ParallelTempering
    from main.precision_models import runtime, temperatures
    L, R, seed, interval, n = 128, 64, 170, 2, 4
    temps = temperatures("he3", R)
    rt = runtime("he3", L, temps, False, seed)
    ptg = ParallelTempering(rt, exchange_interval=interval, seed=seed + 1)
    for _ in range(n):
        ptg.sweep()
    model = rt.model
    state = model.initialize_state(L, R, seed).copy()
    rng = oracle.rng_init(R * L * L, seed)
    pt = _PT(R, (1.0 / temps).astype(np.float32), seed + 1)
    counts = np.zeros((3, R, L, L), np.int64)
    for sweep in range(1, n + 1):
        for color in range(4):
            oracle.he3_gaussian_color(state, rng, pt.betas, color, model.betas, model.K, rt.gaussian.rho,
                                      rt.gaussian.mass, counts, model.g_d)
        if sweep % interval == 0:
            pt.exchange(oracle.he3_energy(state, model.betas, model.K, model.g_d), sweep)
    assert pt.counts[1].sum() > 0
    np.testing.assert_array_equal(rt.snapshot(), state)
    np.testing.assert_array_equal(device_rng_as_pairs(rt.rng_states), rng)
    np.testing.assert_array_equal(ptg.walker_to_temperature.copy_to_host(), pt.w2t)
    diag = rt.sampling_diagnostics()
    np.testing.assert_array_equal(np.asarray(diag["attempts"]), counts[0].sum(axis=(1, 2)))
    np.testing.assert_array_equal(np.asarray(diag["accepted"]), counts[1].sum(axis=(1, 2)))
    np.testing.assert_array_equal(np.asarray(diag["nonfinite_rejections"]), counts[2].sum(axis=(1, 2)))


def test_he3_gaussian_complex128_100_sweeps_within_1e12(gpu):
    from models.he3.model import He3Model
    from models.he3.sampling import He3GaussianRuntime
    model = He3Model(pressure=20, g_d=0.001, dtype=np.complex128)
    L, R, seed = 8, 3, 311
    rt = He3GaussianRuntime(model, L, [0.3, 0.4, 0.6], seed=seed, rho=-0.9)
    state = model.initialize_state(L, R, seed).copy()
    rng = oracle.rng_init(R * L * L, seed)
    betas = (1.0 / rt.temperatures).astype(np.float64)
    counts = np.zeros((3, R, L, L), np.int64)
    for _ in range(100):
        rt.sweep()
        for color in range(4):
            oracle.he3_gaussian_color(state, rng, betas, color, model.betas, model.K, rt.gaussian.rho,
                                      rt.gaussian.mass, counts, model.g_d)
    got = rt.snapshot()
    assert np.max(np.abs(got - state)) < 1e-12, np.max(np.abs(got - state))
    diag = rt.sampling_diagnostics()
    assert abs(int(np.sum(diag["accepted"])) - int(counts[1].sum())) <= 2
