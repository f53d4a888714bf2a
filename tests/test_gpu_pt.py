"""GPU replica exchange: the reference's deterministic PT scenarios
(tests/test_parallel_tempering.py, tests/test_pt_walkers.py,
tests/test_energy_cache.py), and the exchange kernel against the oracle on
random energies."""

from unittest.mock import patch

import numpy as np
import pytest

import oracle
from gpu_helpers import device_rng_as_pairs

pytestmark = pytest.mark.gpu


def constant_xy(gpu, angles, exchange_interval=1):
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from models.xy.model import XYModel
    runtime = Runtime(XYModel(J=1.37), L=4, temperatures=np.linspace(0.6, 1.4, len(angles)), seed=31)
    state = np.broadcast_to(np.asarray(angles, dtype=np.float32)[:, None, None],
                            (len(angles), 4, 4)).copy()
    runtime.state.copy_to_device(state)
    pt = ParallelTempering(runtime, exchange_interval=exchange_interval, seed=97)
    return runtime, pt, state


def assert_mappings(pt, temperature_to_walker):
    expected = np.asarray(temperature_to_walker, dtype=np.int32)
    actual = pt.temperature_to_walker.copy_to_host()
    reverse = pt.walker_to_temperature.copy_to_host()
    np.testing.assert_array_equal(actual, expected)
    np.testing.assert_array_equal(reverse[actual], np.arange(pt.R))
    np.testing.assert_array_equal(actual[reverse], np.arange(pt.R))
    ladder_betas = (1.0 / pt.temperatures).astype(np.float32)
    np.testing.assert_array_equal(pt.runtime.betas.copy_to_host(), ladder_betas[reverse])


def xy_reference(state, J):
    spins = np.exp(1j * state.astype(np.float64))
    bonds_x = spins * np.conj(np.roll(spins, -1, axis=1))
    bonds_y = spins * np.conj(np.roll(spins, -1, axis=2))
    return {"energy": -float(J) * np.sum(bonds_x.real + bonds_y.real, axis=(1, 2)),
            "mx": np.sum(spins.real, axis=(1, 2)), "my": np.sum(spins.imag, axis=(1, 2)),
            "twist_cos_x": np.sum(bonds_x.real, axis=(1, 2)),
            "twist_sin_x": np.sum(bonds_x.imag, axis=(1, 2))}


def test_accepted_exchanges_preserve_walkers_and_order_every_channel(gpu):
    runtime, pt, state = constant_xy(gpu, [0.0, 0.5, 1.0, -0.5])
    site_rng = runtime.rng_states.copy_to_host()
    reference = xy_reference(state, runtime.model.J)
    pt.attempt_exchanges()
    order = np.array([1, 0, 3, 2])
    assert_mappings(pt, order)
    np.testing.assert_array_equal(runtime.snapshot(), state)
    np.testing.assert_array_equal(runtime.rng_states.copy_to_host(), site_rng)
    np.testing.assert_array_equal(pt.snapshot(), state[order])
    measured = pt.measure()
    for name, expected in reference.items():
        np.testing.assert_allclose(measured[name], expected[order], rtol=1e-12, atol=1e-10)
    pt.attempt_exchanges()
    assert_mappings(pt, [1, 3, 0, 2])
    d = pt.diagnostics()
    np.testing.assert_array_equal(d["attempts"], [1, 1, 1])
    np.testing.assert_array_equal(d["accepted"], [1, 1, 1])
    np.testing.assert_array_equal(d["acceptance_rates"], [1.0, 1.0, 1.0])


def test_odd_replica_count_and_more_pairs_than_threads(gpu):
    for R in (5, 259, 517):
        _, pt, _ = constant_xy(gpu, np.linspace(-1.0, 1.0, R))
        order = np.arange(R)
        attempts = np.zeros(R - 1, dtype=np.int64)
        for round_index in range(4):
            for t in range(round_index % 2, R - 1, 2):
                order[t:t + 2] = order[t:t + 2][::-1]
                attempts[t] += 1
            pt.attempt_exchanges()
            assert_mappings(pt, order)
            d = pt.diagnostics()
            np.testing.assert_array_equal(d["attempts"], attempts)
            np.testing.assert_array_equal(d["accepted"], attempts)
            assert d["exchange_rounds"] == round_index + 1


def test_nonfinite_energy_prevents_swap_and_diagnostics_raise(gpu):
    runtime, pt, state = constant_xy(gpu, [0.0, 1.0])
    state[0, 0, 0] = np.nan
    runtime.state.copy_to_device(state)
    pt.attempt_exchanges()
    assert np.isnan(pt.energies.copy_to_host()[0])
    assert_mappings(pt, [0, 1])
    with pytest.raises(FloatingPointError):
        pt.diagnostics()


def test_acceptance_sign_and_current_total_ising_energies(gpu):
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from models.ising.model import IsingModel
    runtime = Runtime(IsingModel(J=1.37), L=4, temperatures=[1e-6, 1e6], seed=32)
    pt = ParallelTempering(runtime, exchange_interval=1, seed=98)
    ordered = np.ones((4, 4), dtype=np.int8)
    checkerboard = (1 - 2 * (np.indices((4, 4)).sum(axis=0) % 2)).astype(np.int8)
    magnitude = 2 * float(runtime.model.J) * 4**2
    runtime.state.copy_to_device(np.stack([ordered, checkerboard]))
    pt.attempt_exchanges()
    np.testing.assert_array_equal(pt.energies.copy_to_host(), [-magnitude, magnitude])
    assert_mappings(pt, [0, 1])
    np.testing.assert_array_equal(pt.diagnostics()["accepted"], [0])
    state = np.stack([checkerboard, ordered])
    runtime.state.copy_to_device(state)
    pt.attempt_exchanges()
    np.testing.assert_array_equal(pt.energies.copy_to_host(), [magnitude, -magnitude])
    assert_mappings(pt, [1, 0])
    np.testing.assert_array_equal(pt.diagnostics()["attempts"], [2])
    np.testing.assert_array_equal(pt.diagnostics()["accepted"], [1])
    measured = pt.measure()
    np.testing.assert_array_equal(measured["energy"], [-magnitude, magnitude])
    np.testing.assert_array_equal(measured["magnetization"], [16, 0])
    np.testing.assert_array_equal(runtime.snapshot(), state)


def test_exchange_interval_and_two_walker_round_trip(gpu):
    runtime, pt, _ = constant_xy(gpu, [0.0, 1.0], exchange_interval=2)
    initial = pt.diagnostics()
    assert initial["mean_round_trip_sweeps"] is None
    np.testing.assert_array_equal(initial["walker_round_trips"], [0, 0])
    with patch.object(runtime, "sweep") as sweep:
        for completed in range(1, 7):
            pt.sweep()
            assert pt.sweeps == completed
            assert pt.diagnostics()["exchange_rounds"] == completed // 2
        assert sweep.call_count == 6
    d = pt.diagnostics()
    np.testing.assert_array_equal(d["attempts"], [3])
    np.testing.assert_array_equal(d["accepted"], [3])
    np.testing.assert_array_equal(d["walker_round_trips"], [1, 1])
    np.testing.assert_array_equal(d["walker_round_trip_total_sweeps"], [4, 4])
    assert d["mean_round_trip_sweeps"] == 4.0


def test_interior_walker_round_trip_starts_at_first_endpoint(gpu):
    runtime, pt, _ = constant_xy(gpu, [0.0, 0.5, 1.0])
    with patch.object(runtime, "sweep"):
        for _ in range(6):
            pt.sweep()
        np.testing.assert_array_equal(pt.diagnostics()["walker_round_trips"], [1, 0, 1])
        pt.sweep()
    d = pt.diagnostics()
    np.testing.assert_array_equal(d["walker_round_trips"], [1, 1, 1])
    np.testing.assert_array_equal(d["walker_round_trip_total_sweeps"], [5, 6, 6])


def test_walker_history_chunks_stride_and_manual_exchanges(gpu):
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from models.xy.model import XYModel

    def make(R=4, **kw):
        rt = Runtime(XYModel(), L=4, temperatures=np.linspace(.8, 1.2, R), threads=2, seed=162)
        return rt, ParallelTempering(rt, exchange_interval=1, seed=163, **kw)

    runtime, pt = make(tracked_walkers=range(4), history_stride=1)
    expected = [np.arange(4)]
    with patch.object(runtime, "sweep"):
        for _ in range(150):  # the trail's device buffer grows 64 -> 128 -> 256 rows
            pt.sweep()
            expected.append(pt.walker_to_temperature.copy_to_host())
    history = pt.diagnostics()["walker_history"]
    assert pt._trail.rows.shape[0] == 256 and pt._trail.used == 151
    np.testing.assert_array_equal(history["temperature_indices"], expected)
    np.testing.assert_array_equal(history["sweeps"], np.arange(151))
    np.testing.assert_array_equal(history["exchange_rounds"], np.arange(151))

    runtime, pt = make(tracked_walkers=[3, 0], history_stride=3)
    with patch.object(runtime, "sweep"):
        for _ in range(7):
            pt.sweep()
    history = pt.diagnostics()["walker_history"]
    np.testing.assert_array_equal(history["sweeps"], [0, 3, 6, 7])
    np.testing.assert_array_equal(history["temperature_indices"][-1],
                                  pt.walker_to_temperature.copy_to_host()[[3, 0]])

    _, pt = make(R=2, tracked_walkers=[0, 1], history_stride=1)
    for _ in range(4):
        pt.attempt_exchanges()
    d = pt.diagnostics()
    np.testing.assert_array_equal(d["walker_history"]["temperature_indices"],
                                  [[0, 1], [1, 0], [0, 1], [1, 0], [0, 1]])
    np.testing.assert_array_equal(d["walker_round_trips"], [2, 2])


@pytest.mark.parametrize("beta_dtype", [np.float32, np.float64])
def test_exchange_kernel_matches_oracle_on_random_energies(gpu, beta_dtype):
    from mcgpu import _lib, as_pointer, create_xoroshiro128p_states, dtype_code
    rng = np.random.default_rng(5)
    R = 37
    betas = (1.0 / np.linspace(0.5, 2.0, R)).astype(beta_dtype)
    w2t = np.arange(R, dtype=np.int32)
    t2w = np.arange(R, dtype=np.int32)
    counts = np.zeros((3, R - 1), dtype=np.int64)
    trips = np.zeros((5, R), dtype=np.int64)
    trips[0] = -1
    trips[2] = -1
    trips[0, 0], trips[0, -1] = 0, R - 1
    trips[2, [0, -1]] = 0
    d_rng = create_xoroshiro128p_states(R - 1, 99)
    h_rng = oracle.rng_init(R - 1, 99)
    dev = {name: gpu.to_device(a) for name, a in
           dict(betas=betas, w2t=w2t, t2w=t2w, counts=counts, trips=trips).items()}
    for step in range(60):
        energies = rng.normal(scale=3.0, size=R)
        if step % 17 == 5:
            energies[rng.integers(R)] = np.inf
        d_e = gpu.to_device(energies)
        parity = step % 2
        _lib.call("mc_pt_exchange", None, dtype_code(beta_dtype), as_pointer(d_e),
                  as_pointer(dev["betas"]), as_pointer(dev["w2t"]), as_pointer(dev["t2w"]),
                  as_pointer(d_rng), as_pointer(dev["counts"]), as_pointer(dev["trips"]), R, parity,
                  step)
        oracle.pt_exchange(energies, betas, w2t, t2w, h_rng, counts, trips, parity, step)
    for name, host in dict(betas=betas, w2t=w2t, t2w=t2w, counts=counts, trips=trips).items():
        np.testing.assert_array_equal(dev[name].copy_to_host(), host, err_msg=name)
    np.testing.assert_array_equal(device_rng_as_pairs(d_rng), h_rng)
    assert counts[1].sum() > 0 and counts[2].sum() > 0


def test_invalid_ladders_and_exchange_intervals(gpu):
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from models.ising.model import IsingModel
    for temperatures in ([1.0], [1.0, 1.0], [2.0, 1.0], [1.0, 2.0, 1.5]):
        runtime = Runtime(IsingModel(), L=4, temperatures=temperatures)
        with pytest.raises(ValueError):
            ParallelTempering(runtime)
    runtime = Runtime(IsingModel(), L=4, temperatures=[1.0, 1.0 + 1e-10])
    with pytest.raises(ValueError):
        ParallelTempering(runtime)
    runtime = Runtime(IsingModel(), L=4, temperatures=[1.0, 2.0])
    for interval in (0, -1, 1.5, True):
        with pytest.raises(ValueError):
            ParallelTempering(runtime, exchange_interval=interval)


# ---------------------------------------------------------------- energy cache

def _he3_runtime(L=6, replicas=3):
    from core.runtime import Runtime
    from models.he3.model import He3Model
    runtime = Runtime(He3Model(pressure=20), L=L, temperatures=np.linspace(0.2, 0.5, replicas),
                      threads=2, seed=211)
    rng = np.random.default_rng(212)
    state = 0.25 * (rng.normal(size=(replicas, L, L, 3, 2)) + 1j * rng.normal(size=(replicas, L, L, 3, 2)))
    runtime.state.copy_to_device(state.astype(np.complex64))
    runtime.invalidate_energy_cache()
    return runtime


def _cache_energy(gpu, runtime):
    energies = gpu.device_array(runtime.R, np.float64)
    runtime.measure_energy(energies)
    return energies


def test_cached_measurements_and_invalidation(gpu):
    runtime = _he3_runtime()
    expected = runtime.model.measure(runtime.state)
    energies = _cache_energy(gpu, runtime)
    original = energies.copy_to_host()
    with patch.object(runtime.model, "measure_with_energy",
                      wraps=runtime.model.measure_with_energy) as cached:
        actual = runtime.measure()
    cached.assert_called_once()
    for name in actual:
        np.testing.assert_allclose(actual[name], expected[name], rtol=2e-6, atol=1e-5)
    np.testing.assert_array_equal(actual["energy"], original)

    runtime.betas.copy_to_device(np.zeros(runtime.R, dtype=np.float32))
    runtime.sweep()
    with patch.object(runtime.model, "measure_with_energy",
                      wraps=runtime.model.measure_with_energy) as cached:
        actual = runtime.measure()
    cached.assert_not_called()
    assert not np.allclose(actual["energy"], original)


def test_failed_sweep_and_failed_refresh_never_publish_stale_energy(gpu):
    runtime = _he3_runtime(L=4, replicas=2)
    _cache_energy(gpu, runtime)
    changed = (1.2 * runtime.snapshot()).astype(np.complex64)

    def fail_after_mutation(*args, **kwargs):
        runtime.state.copy_to_device(changed)
        raise RuntimeError("interrupted update")

    with patch.object(runtime.model, "update_color", side_effect=fail_after_mutation):
        with pytest.raises(RuntimeError, match="interrupted update"):
            runtime.sweep()
    with patch.object(runtime.model, "measure_with_energy",
                      wraps=runtime.model.measure_with_energy) as cached:
        runtime.measure()
    cached.assert_not_called()

    energies = _cache_energy(gpu, runtime)

    def fail_after_overwrite(state, output):
        output.copy_to_device(np.full(runtime.R, 12345.0, dtype=np.float64))
        raise RuntimeError("interrupted energy calculation")

    with patch.object(runtime.model, "measure_energy", side_effect=fail_after_overwrite):
        with pytest.raises(RuntimeError, match="interrupted energy calculation"):
            runtime.measure_energy(energies)
    with patch.object(runtime.model, "measure_with_energy",
                      wraps=runtime.model.measure_with_energy) as cached:
        runtime.measure()
    cached.assert_not_called()


def test_ising_xy_kawamura_reuse_energy_and_ordering(gpu):
    from core.runtime import Runtime
    from models.ising.model import IsingModel
    from models.kawamura.model import KawamuraModel
    from models.xy.model import XYModel
    rng = np.random.default_rng(215)
    for model in (XYModel(J=1.3), IsingModel(J=0.9), KawamuraModel(J=1.1)):
        for L in (4, 6, 20):
            runtime = Runtime(model, L=L, temperatures=[0.4, 0.6, 0.8], threads=2, seed=216)
            if isinstance(model, XYModel):
                state = rng.uniform(-np.pi, np.pi, (3, L, L)).astype(np.float32)
            elif isinstance(model, IsingModel):
                state = rng.choice([-1, 1], size=(3, L, L)).astype(np.int8)
            else:
                state = np.empty((3, L, L, 5), dtype=np.float32)
                state[..., 0] = rng.uniform(-np.pi, np.pi, (3, L, L))
                state[..., 1] = rng.choice([-1, 1], size=(3, L, L))
                spin = rng.normal(size=(3, L, L, 3))
                state[..., 2:] = spin / np.linalg.norm(spin, axis=-1, keepdims=True)
            runtime.state.copy_to_device(state)
            runtime.invalidate_energy_cache()
            expected = model.measure(runtime.state)
            energies = _cache_energy(gpu, runtime)
            original = energies.copy_to_host()
            actual = runtime.measure()
            for name in expected:
                np.testing.assert_allclose(actual[name], expected[name], rtol=1e-12, atol=1e-10)
            np.testing.assert_array_equal(actual["energy"], original)
            order = np.array([2, 0, 1], dtype=np.int32)
            mapping = gpu.to_device(order)
            sentinel = np.array([12345., -6789., .125])
            energies.copy_to_device(sentinel)
            reused = runtime.measure(temperature_indices=mapping)
            np.testing.assert_array_equal(reused["energy"][order], sentinel)
            for name in expected.keys() - {"energy"}:
                np.testing.assert_allclose(reused[name][order], expected[name], rtol=1e-12, atol=1e-10)
            runtime.sweep()
            fresh = runtime.measure()
            reference = model.measure(runtime.state)
            for name in reference:
                np.testing.assert_array_equal(fresh[name], reference[name])


def test_simulation_with_pt_schedule_and_samples(gpu):
    from core.simulation import Simulation
    runtime, pt, state = constant_xy(gpu, [0.0, 1.0], exchange_interval=3)
    simulation = Simulation(pt, thermalization_sweeps=4, measurement_sweeps=7, sample_interval=2)
    with patch.object(runtime, "sweep") as sweep:
        run = simulation.run()
    assert sweep.call_count == 11
    np.testing.assert_array_equal(run["metadata"]["sample_sweeps"], [6, 8, 10])
    mx = xy_reference(state, runtime.model.J)["mx"]
    np.testing.assert_allclose(run["samples"]["mx"], [mx, mx, mx[::-1]], atol=1e-12)
    d = run["metadata"]["parallel_tempering"]
    assert d["exchange_interval"] == 3 and d["exchange_seed"] == 97
    assert d["sweeps"] == 11 and d["exchange_rounds"] == 3
    np.testing.assert_array_equal(d["walker_round_trips"], [1, 1])
    assert d["mean_round_trip_sweeps"] == 6.0
