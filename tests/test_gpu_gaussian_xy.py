"""GPU parity of the Gaussian-proposal XY updates.

* plain path (Runtime + XYModel(proposal="gaussian")): float32 trajectories,
  RNG streams and per-temperature acceptance counters bit-identical to the
  oracle, float64 within 1e-12; with parallel tempering the widths follow
  the temperatures (oracle loop with the same exchanges);
* batched engine (float32 fast restatement): trajectories, streams and
  counters bit-identical to the oracle's fast Gaussian update, with one and
  several strips per replica, with and without exchanges;
* width adaptation reaches the target acceptance and freezes afterwards.
"""

import numpy as np
import pytest

import oracle
from gpu_helpers import device_rng_as_pairs

pytestmark = pytest.mark.gpu


def _plain_reference(theta, rng, betas, table, sigmas, nsweeps, acc):
    for _ in range(nsweeps):
        for color in range(4):
            oracle.xy_gauss_color(theta, rng, betas, color, 1.0, table, sigmas, acc)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_plain_gaussian_trajectory_and_counters(gpu, dtype):
    from core.runtime import Runtime
    from models.xy.model import XYModel
    L, R, seed = 16, 3, 31
    temps = np.array([0.7, 0.9, 1.3])
    sig = np.array([0.4, 1.0, 2.9])
    model = XYModel(dtype=dtype, start="random", proposal="gaussian", sigma=sig)
    rt = Runtime(model, L, temps, seed=seed)
    n = 12 if dtype == np.float32 else 100
    for _ in range(n):
        rt.sweep()
    theta = model.initialize_state(L, R, seed).copy()
    rng = oracle.rng_init(R * L * L, seed)
    betas = (1.0 / temps).astype(dtype)
    acc = np.zeros(R, np.int64)
    _plain_reference(theta, rng, betas, betas.copy(), sig.astype(dtype), n, acc)
    tables = model.proposal_tables(rt)
    got_acc = tables.accepted.copy_to_host().astype(np.int64)
    if dtype == np.float32:
        np.testing.assert_array_equal(rt.snapshot(), theta)
        np.testing.assert_array_equal(device_rng_as_pairs(rt.rng_states), rng)
        np.testing.assert_array_equal(got_acc, acc)
    else:
        d = np.angle(np.exp(1j * (rt.snapshot() - theta)))
        assert np.abs(d).max() < 1e-12
        assert np.abs(got_acc - acc).max() <= 2
    np.testing.assert_allclose(tables.acceptance(), got_acc / (n * L * L))


def test_plain_gaussian_widths_follow_temperatures_through_exchanges(gpu):
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from models.xy.model import XYModel
    L, R, seed, interval = 16, 4, 9, 2
    temps = np.geomspace(0.7, 1.1, R)
    sig = np.array([0.3, 0.8, 1.6, 2.4])
    model = XYModel(start="random", proposal="gaussian", sigma=sig)
    rt = Runtime(model, L, temps, seed=seed)
    pt = ParallelTempering(rt, exchange_interval=interval, seed=seed + 1)
    nsweeps = 20
    for _ in range(nsweeps):
        pt.sweep()
    theta = model.initialize_state(L, R, seed).copy()
    rng = oracle.rng_init(R * L * L, seed)
    table = (1.0 / temps).astype(np.float32)
    betas = table.copy()
    acc = np.zeros(R, np.int64)
    w2t = np.arange(R, dtype=np.int32)
    t2w = np.arange(R, dtype=np.int32)
    counts = np.zeros((3, R - 1), np.int64)
    trips = np.zeros((5, R), np.int64)
    trips[0] = trips[2] = -1
    trips[0, 0], trips[0, -1] = 0, R - 1
    trips[2, [0, -1]] = 0
    pt_rng = oracle.rng_init(R - 1, seed + 1)
    parity = 0
    for sweep in range(1, nsweeps + 1):
        _plain_reference(theta, rng, betas, table, sig.astype(np.float32), 1, acc)
        if sweep % interval == 0:
            e = oracle.xy_energy(theta, 1.0)
            oracle.pt_exchange(e, betas, w2t, t2w, pt_rng, counts, trips, parity, sweep)
            parity = 1 - parity
    np.testing.assert_array_equal(rt.snapshot(), theta)
    np.testing.assert_array_equal(rt.betas.copy_to_host(), betas)
    assert counts[1].sum() > 0  # some exchanges happened
    np.testing.assert_array_equal(model.proposal_tables(rt).accepted.copy_to_host().astype(np.int64), acc)


def _fast_reference(theta, c, s, rng, betas, sig_walker, nsweeps, acc_walker):
    for _ in range(nsweeps):
        for color in range(4):
            oracle.xyfast_gauss_color(theta, c, s, rng, betas, sig_walker, color, 1.0, acc_walker)


@pytest.mark.parametrize("L,R,H,threads", [(16, 3, 16, 256), (32, 4, 8, 256), (64, 2, 4, 256),
                                           (128, 2, 16, 512), (128, 2, 32, 512)])
def test_batched_gaussian_trajectory_matches_oracle(gpu, L, R, H, threads):
    from core.batched import BatchedRun
    from models.xy.model import XYModel
    temps = np.geomspace(0.7, 1.1, R)
    sig = np.linspace(0.5, 2.5, R)
    model = XYModel(start="random", proposal="gaussian", sigma=sig)
    br = BatchedRun(model, L, temps, seed=77, exchange_interval=0, strip_rows=H, threads=threads)
    br.run(16)
    theta = model.initialize_state(L, R, 77).copy()
    c, s = oracle.xyfast_cache(theta)
    rng = oracle.rng_init(R * L * L, 77)
    betas = (1.0 / temps).astype(np.float32)
    acc = np.zeros(R, np.int64)
    _fast_reference(theta, c, s, rng, betas, sig.astype(np.float32), 16, acc)
    np.testing.assert_array_equal(br.snapshot(), theta)
    np.testing.assert_array_equal(device_rng_as_pairs(br.rng_states), rng)
    np.testing.assert_array_equal(br._accepted.copy_to_host().astype(np.int64), acc)
    np.testing.assert_allclose(br.acceptance_by_temperature(), acc / (16 * L * L))


def test_batched_gaussian_with_exchanges_matches_oracle_loop(gpu):
    from core.batched import BatchedRun
    from models.xy.model import XYModel
    L, R, interval = 32, 6, 5
    temps = np.geomspace(0.7, 1.1, R)
    sig = np.linspace(0.4, 2.8, R)
    model = XYModel(start="random", proposal="gaussian", sigma=sig)
    br = BatchedRun(model, L, temps, seed=13, exchange_interval=interval, pt_seed=14, strip_rows=8)
    br.run(40)
    theta = model.initialize_state(L, R, 13).copy()
    c, s = oracle.xyfast_cache(theta)
    rng = oracle.rng_init(R * L * L, 13)
    betas = (1.0 / temps).astype(np.float32)
    w2t = np.arange(R, dtype=np.int32)
    t2w = np.arange(R, dtype=np.int32)
    counts = np.zeros((3, R - 1), np.int64)
    trips = np.zeros((5, R), np.int64)
    trips[0] = trips[2] = -1
    trips[0, 0], trips[0, -1] = 0, R - 1
    trips[2, [0, -1]] = 0
    pt_rng = oracle.rng_init(R - 1, 14)
    acc_t = np.zeros(R, np.int64)
    parity = 0
    for sweep in range(1, 41):
        acc_w = np.zeros(R, np.int64)
        _fast_reference(theta, c, s, rng, betas, sig.astype(np.float32)[w2t], 1, acc_w)
        np.add.at(acc_t, w2t, acc_w)
        if sweep % interval == 0:
            e = oracle.xyfast_observables(c, s, 1.0)[0]
            oracle.pt_exchange(e, betas, w2t, t2w, pt_rng, counts, trips, parity, sweep)
            parity = 1 - parity
    assert counts[1].sum() > 0
    np.testing.assert_array_equal(br.snapshot(), theta)
    np.testing.assert_array_equal(br.betas.copy_to_host(), betas)
    np.testing.assert_array_equal(br._accepted.copy_to_host().astype(np.int64), acc_t)


def test_width_adaptation_reaches_target_and_freezes(gpu):
    from core.batched import BatchedRun
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from models.xy.model import XYModel
    from models.xy.proposals import SIGMA_MAX, adapt
    L, R = 32, 6
    temps = np.geomspace(0.7, 1.1, R)
    br = BatchedRun(XYModel(start="random", proposal="gaussian", sigma=1.0), L, temps, seed=5,
                    exchange_interval=10, strip_rows=8)
    br.adapt_sigmas(1500, interval=50)
    frozen = br._sigmas_host.copy()
    br.run(1000)
    rates = br.acceptance_by_temperature()
    np.testing.assert_array_equal(br._sigmas_host, frozen)
    inner = frozen < SIGMA_MAX * 0.999
    assert inner.any()
    assert np.all(np.abs(rates[inner] - 0.5) < 0.05), (rates, frozen)

    model = XYModel(start="random", proposal="gaussian", sigma=1.0)
    rt = Runtime(model, 16, temps, seed=6)
    pt = ParallelTempering(rt, exchange_interval=10, seed=7)
    tables = model.proposal_tables(rt)
    adapt(pt, tables, 800, interval=50)
    for _ in range(400):
        pt.sweep()
    rates = tables.acceptance()
    inner = tables.sigmas_host < SIGMA_MAX * 0.999
    assert np.all(np.abs(rates[inner] - 0.5) < 0.06), (rates, tables.sigmas_host)
