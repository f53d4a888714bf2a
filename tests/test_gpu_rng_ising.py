"""GPU parity: RNG stream initialisation and the Ising path, bit-exact against
the CPU oracle, plus the reference's Ising behavioural tests
(tests/test_simulation.py:72-135 of the reference)."""

import numpy as np
import pytest

import oracle
from gpu_helpers import device_rng_as_pairs, oracle_rng

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,seed,sub", [
    (1, 42, 0), (1024, 74, 0), (65537, 0, 0), (513, 2**63 + 5, 0),
    (37, -1, 0), (100, 7, 3), (4096, 170, 0),
])
def test_rng_streams_match_oracle(gpu, n, seed, sub):
    dev = gpu.create_xoroshiro128p_states(n, seed, sub)
    assert dev.shape == (n,)
    np.testing.assert_array_equal(device_rng_as_pairs(dev), oracle.rng_init(n, seed, sub))


@pytest.mark.parametrize("L,R,start,seed", [
    (4, 2, "cold", 42), (6, 3, "random", 7), (20, 4, "random", 11),
    (32, 8, "random", 74), (64, 5, "cold", 106),
])
def test_ising_sweeps_bit_identical_to_oracle(gpu, L, R, start, seed):
    from core.runtime import Runtime
    from models.ising.model import IsingModel

    temperatures = np.linspace(2.0, 2.6, R)
    model = IsingModel(J=1.0, start=start)
    rt = Runtime(model, L, temperatures, seed=seed)
    state = model.initialize_state(L, R, seed)
    rng = oracle_rng(rt)
    betas = (1.0 / temperatures).astype(np.float32)
    for sweep in range(25):
        rt.sweep()
        oracle.ising_sweeps(state, rng, betas, model.J, 1)
        if sweep in (0, 24):
            np.testing.assert_array_equal(rt.snapshot(), state, err_msg=f"sweep {sweep}")
    np.testing.assert_array_equal(device_rng_as_pairs(rt.rng_states), rng)
    measured = rt.measure()
    ref = oracle.ising_measure(state, model.J)
    np.testing.assert_array_equal(measured["energy"], ref[0])
    np.testing.assert_array_equal(measured["magnetization"], ref[1])


def test_ising_nonunit_coupling_and_per_colour_parity(gpu):
    from core.runtime import Runtime
    from models.ising.model import IsingModel

    L, R = 10, 3
    model = IsingModel(J=1.37, start="random")
    rt = Runtime(model, L, [0.9, 1.7, 3.3], seed=5)
    state = model.initialize_state(L, R, 5)
    rng = oracle_rng(rt)
    betas = (1.0 / rt.temperatures).astype(np.float32)
    for color in (0, 1, 2, 3, 3, 1, 0, 2):
        model.update_color(rt.state, rt.rng_states, rt.betas, color, rt.blocks_per_grid,
                           rt.threads_per_block)
        oracle.ising_color(state, rng, betas, color, model.J)
        np.testing.assert_array_equal(rt.snapshot(), state, err_msg=f"color {color}")


def test_ordered_and_alternating_ising_measurements(gpu):
    from models.ising.model import IsingModel

    model = IsingModel(J=1.37)
    L = 6
    checkerboard = 1 - 2 * (np.indices((L, L)).sum(axis=0) % 2)
    spins = np.stack((np.ones((L, L)), -np.ones((L, L)), checkerboard))
    measured = model.measure(gpu.to_device(spins.astype(np.int8)))
    expected = {
        "energy": float(model.J) * L**2 * np.array([-2.0, -2.0, 2.0]),
        "magnetization": np.array([L**2, -L**2, 0]),
    }
    assert measured.keys() == expected.keys()
    for name, values in expected.items():
        assert measured[name].dtype == np.float64
        np.testing.assert_array_equal(measured[name], values)


def test_ising_measurements_match_cpu_bond_count(gpu):
    from models.ising.model import IsingModel

    model = IsingModel(J=1.37)
    rng = np.random.default_rng(18)
    for L in (4, 6, 32, 64, 256):
        spins = rng.choice(np.array([-1, 1], dtype=np.int8), (3, L, L))
        measured = model.measure(gpu.to_device(spins))
        equal_bonds = (np.sum(spins == np.roll(spins, -1, axis=1), axis=(1, 2))
                       + np.sum(spins == np.roll(spins, -1, axis=2), axis=(1, 2)))
        energy = float(model.J) * (2 * L**2 - 2 * equal_bonds)
        np.testing.assert_array_equal(measured["energy"], energy)
        np.testing.assert_array_equal(measured["magnetization"], np.sum(spins, axis=(1, 2)))
        out = gpu.device_array(3, np.float64)
        model.measure_energy(gpu.to_device(spins), out)
        np.testing.assert_array_equal(out.copy_to_host(), energy)


def test_ising_updates_in_temperature_limits(gpu):
    from core.runtime import Runtime
    from models.ising.model import IsingModel

    L = 6
    runtime = Runtime(IsingModel(J=1.37), L, [1.0, 2.0], threads=2)
    rng = np.random.default_rng(19)
    expected = rng.choice(np.array([-1, 1], dtype=np.int8), (2, L, L))
    runtime.state.copy_to_device(expected)
    # beta = 1e8 rejects every positive-energy move; beta = 0 accepts all.
    runtime.betas.copy_to_device(np.array([1e8, 0.0], dtype=np.float32))
    for color in range(4):
        for r in range(2):
            for i in range(color // 2, L, 2):
                for j in range(color % 2, L, 2):
                    neighbors = (int(expected[r, (i - 1) % L, j]) + int(expected[r, (i + 1) % L, j])
                                 + int(expected[r, i, (j - 1) % L]) + int(expected[r, i, (j + 1) % L]))
                    if r == 1 or expected[r, i, j] * neighbors <= 0:
                        expected[r, i, j] *= -1
        runtime.model.update_color(runtime.state, runtime.rng_states, runtime.betas, color,
                                   runtime.blocks_per_grid, runtime.threads_per_block)
        np.testing.assert_array_equal(runtime.snapshot(), expected)


def test_scattered_and_cached_measurements(gpu):
    from models.ising.model import IsingModel

    model = IsingModel(J=0.9)
    rng = np.random.default_rng(215)
    spins = rng.choice([-1, 1], size=(3, 20, 20)).astype(np.int8)
    dev = gpu.to_device(spins)
    ref = oracle.ising_measure(spins, model.J)
    order = np.array([2, 0, 1], dtype=np.int32)
    ordered = model.measure(dev, gpu.to_device(order))
    np.testing.assert_array_equal(ordered["energy"][order], ref[0])
    sentinel = np.array([12345.0, -6789.0, 0.125])
    cached = model.measure(dev, gpu.to_device(order), gpu.to_device(sentinel))
    np.testing.assert_array_equal(cached["energy"][order], sentinel)
    np.testing.assert_array_equal(cached["magnetization"][order], ref[1])
