"""GPU parity for XY, Kawamura and He3: trajectories against the CPU oracle
(bit-identical in float32, within 1e-12 per site in float64 over 100 sweeps),
measurements against the oracle and the reference's closed forms
(tests/test_simulation.py:25-70, tests/test_kawamura.py, tests/test_he3.py)."""

import numpy as np
import pytest

import oracle
from gpu_helpers import device_rng_as_pairs, oracle_color, oracle_rng

pytestmark = pytest.mark.gpu


def _runtime(model, L, temperatures, seed):
    from core.runtime import Runtime
    return Runtime(model, L, temperatures, seed=seed)


def _trajectory_parity(kind, model, L, R, seed, nsweeps, temperatures, exact, check_every=None):
    rt = _runtime(model, L, temperatures, seed)
    state = model.initialize_state(L, R, seed).copy()
    rng = oracle_rng(rt)
    betas = (1.0 / rt.temperatures).astype(model.beta_dtype)
    for sweep in range(nsweeps):
        rt.sweep()
        for color in range(4):
            oracle_color(kind, state, rng, betas, model, color)
        if check_every and (sweep + 1) % check_every == 0:
            got = rt.snapshot()
            if exact:
                np.testing.assert_array_equal(got, state, err_msg=f"sweep {sweep}")
    got = rt.snapshot()
    if exact:
        np.testing.assert_array_equal(got, state)
        np.testing.assert_array_equal(device_rng_as_pairs(rt.rng_states), rng)
    else:
        np.testing.assert_allclose(got, state, rtol=0, atol=1e-12)
        np.testing.assert_array_equal(device_rng_as_pairs(rt.rng_states), rng)
    return rt, state


# ----------------------------------------------------------------------- XY

@pytest.mark.parametrize("L,R,start,seed", [(4, 2, "ordered", 3), (16, 4, "random", 170),
                                            (34, 3, "random", 9)])
def test_xy_float32_trajectory_bit_identical(gpu, L, R, start, seed):
    from models.xy.model import XYModel
    model = XYModel(J=1.0, delta=1.5, start=start)
    temps = np.geomspace(0.7, 1.1, R)
    rt, state = _trajectory_parity("xy", model, L, R, seed, 40, temps, exact=True, check_every=10)
    ref = oracle.xy_measure(state, model.J)
    got = rt.measure()
    for c, name in enumerate(("energy", "mx", "my", "twist_cos_x", "twist_sin_x")):
        np.testing.assert_allclose(got[name], ref[c], rtol=1e-12, atol=1e-10)


def test_xy_float64_trajectory_within_1e12(gpu):
    from models.xy.model import XYModel
    model = XYModel(J=1.0, delta=1.5, dtype=np.float64, start="random")
    _trajectory_parity("xy", model, 16, 4, 21, 100, np.geomspace(0.7, 1.1, 4), exact=False)


def test_ordered_xy_measurements_include_coupling(gpu):
    from models.xy.model import XYModel
    model = XYModel(J=1.37)
    measured = _runtime(model, 4, [0.8, 1.0], 42).measure()
    expected = {"energy": -2 * model.J * 16, "mx": 16, "my": 0, "twist_cos_x": 16, "twist_sin_x": 0}
    assert measured.keys() == expected.keys()
    for name, value in expected.items():
        assert measured[name].shape == (2,)
        assert measured[name].dtype == np.float64
        np.testing.assert_allclose(measured[name], value, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_xy_measurements_match_independent_cpu_reference(gpu, dtype):
    from models.xy.model import XYModel
    model = XYModel(J=1.37, dtype=dtype)
    rng = np.random.default_rng(17)
    for L in (4, 6, 32, 128):
        angles = rng.uniform(-np.pi, np.pi, (3, L, L)).astype(dtype)
        dev = gpu.to_device(angles)
        measured = model.measure(dev)
        spins = np.exp(1j * angles.astype(np.float64))
        bonds_x = spins * np.conj(np.roll(spins, -1, axis=1))
        bonds_y = spins * np.conj(np.roll(spins, -1, axis=2))
        expected = {
            "energy": -float(model.J) * np.sum(bonds_x.real + bonds_y.real, axis=(1, 2)),
            "mx": np.sum(spins.real, axis=(1, 2)), "my": np.sum(spins.imag, axis=(1, 2)),
            "twist_cos_x": np.sum(bonds_x.real, axis=(1, 2)),
            "twist_sin_x": np.sum(bonds_x.imag, axis=(1, 2)),
        }
        for name in expected:
            np.testing.assert_allclose(measured[name], expected[name], rtol=1e-12, atol=1e-10)
        energies = gpu.device_array(3, np.float64)
        model.measure_energy(dev, energies)
        np.testing.assert_array_equal(energies.copy_to_host(), measured["energy"])


# ----------------------------------------------------------------- Kawamura

def _kawamura_reference(state, J):
    theta, tau, d = state[..., 0], state[..., 1], state[..., 2:]
    phase = np.exp(1j * theta)
    bonds = []
    for axis in (1, 2):
        weight = (1 + tau * np.roll(tau, -1, axis)) * np.sum(d * np.roll(d, -1, axis), axis=-1)
        bonds.append(weight * phase * np.conj(np.roll(phase, -1, axis)))
    tensor = np.sum(d[..., :, None] * d[..., None, :] - np.eye(3) / 3, axis=(1, 2))
    return {
        "energy": -J * np.sum(bonds[0].real + bonds[1].real, axis=(1, 2)),
        "chirality": tau.sum(axis=(1, 2)),
        "px": (phase**2).real.sum(axis=(1, 2)), "py": (phase**2).imag.sum(axis=(1, 2)),
        "spin2": np.sum(tensor**2, axis=(1, 2)) / 3,
        "twist_cos_x": bonds[0].real.sum(axis=(1, 2)),
        "twist_sin_x": bonds[0].imag.sum(axis=(1, 2)),
    }


def _kawamura_random_state(L, R=3, seed=63, dtype=np.float32):
    rng = np.random.default_rng(seed)
    state = np.empty((R, L, L, 5), dtype=dtype)
    state[..., 0] = rng.uniform(-np.pi, np.pi, state.shape[:3])
    state[..., 1] = rng.choice([-1, 1], state.shape[:3])
    d = rng.normal(size=(*state.shape[:3], 3))
    state[..., 2:] = d / np.linalg.norm(d, axis=-1, keepdims=True)
    return state


@pytest.mark.parametrize("start", ["ordered", "random"])
def test_kawamura_float32_trajectory_bit_identical(gpu, start):
    from models.kawamura.model import KawamuraModel
    model = KawamuraModel(J=1.0, delta=1.5, start=start)
    _trajectory_parity("kawamura", model, 12, 3, 31, 20, [0.65, 0.72, 0.8], exact=True,
                       check_every=5)


def test_kawamura_float64_trajectory_within_1e12(gpu):
    from models.kawamura.model import KawamuraModel
    model = KawamuraModel(J=1.0, delta=1.5, dtype=np.float64, start="random")
    _trajectory_parity("kawamura", model, 12, 3, 32, 100, [0.65, 0.72, 0.8], exact=False)


def test_kawamura_measurements_energy_and_order(gpu):
    from models.kawamura.model import KawamuraModel
    model = KawamuraModel(J=1.37)
    for L in (4, 6, 20):
        state = _kawamura_random_state(L)
        dev = gpu.to_device(state)
        expected = _kawamura_reference(state.astype(float), float(model.J))
        measured = model.measure(dev)
        for name in expected:
            np.testing.assert_allclose(measured[name], expected[name], atol=1e-10, rtol=1e-12)
        energy = gpu.device_array(3, np.float64)
        model.measure_energy(dev, energy)
        np.testing.assert_allclose(energy.copy_to_host(), expected["energy"], atol=1e-10)
        order = np.array([2, 0, 1], dtype=np.int32)
        ordered = model.measure(dev, gpu.to_device(order))
        for name in expected:
            np.testing.assert_allclose(ordered[name][order], expected[name], atol=1e-10)


def test_kawamura_ordered_closed_forms_and_limits(gpu):
    from models.kawamura.model import KawamuraModel
    model = KawamuraModel(J=1.37)
    rt = _runtime(model, 4, [0.6, 0.9], 42)
    r = rt.measure()
    expected = {"energy": -4 * float(model.J) * 16, "chirality": 16, "px": 16, "py": 0,
                "spin2": 2 * 16**2 / 9, "twist_cos_x": 32, "twist_sin_x": 0}
    for name, value in expected.items():
        np.testing.assert_allclose(r[name], value, atol=1e-12)
    # beta = 0 accepts every move (u < exp(0) = 1), so every tau flips once per
    # sweep: after an odd number of sweeps all chiralities are reversed.  (The
    # reference test asks for both signs after 20 sweeps, which this exact
    # limit rules out.)
    hot = _runtime(KawamuraModel(), 6, [0.6, 1.0], 42)
    hot.betas.copy_to_device(np.zeros(2, dtype=np.float32))
    for _ in range(21):
        hot.sweep()
    state = hot.snapshot()
    assert np.all(np.abs(state[..., 0]) <= np.pi + 1e-6)
    np.testing.assert_array_equal(state[..., 1], -1.0)
    np.testing.assert_allclose(np.linalg.norm(state[..., 2:], axis=-1), 1, atol=2e-6)
    assert np.std(state[..., 0]) > 0.1 and np.std(state[..., 2:]) > 0.1
    cold = _runtime(KawamuraModel(), 4, [0.6, 1.0], 42)
    cold.betas.copy_to_device(np.full(2, 1e8, dtype=np.float32))
    for _ in range(4):
        cold.sweep()
    np.testing.assert_allclose(cold.measure()["energy"], -64, atol=1e-5)


# ---------------------------------------------------------------------- He3

def _he3_random_state(L=6, R=3, seed=253, dtype=np.complex64):
    rng = np.random.default_rng(seed)
    shape = (R, L, L, 3, 2)
    return (0.3 * (rng.normal(size=shape) + 1j * rng.normal(size=shape))).astype(dtype)


def test_he3_cartesian_float32_trajectory_bit_identical(gpu):
    from models.he3.model import He3Model
    model = He3Model(pressure=20, g_d=0.001)
    rt = _runtime(model, 8, [0.3, 0.4, 0.6], 227)
    state = _he3_random_state(8, 3, 5)
    rt.state.copy_to_device(state)
    rt.invalidate_energy_cache()
    rng = oracle_rng(rt)
    betas = (1.0 / rt.temperatures).astype(np.float32)
    for _ in range(15):
        rt.sweep()
        for color in range(4):
            oracle_color("he3", state, rng, betas, model, color)
    np.testing.assert_array_equal(rt.snapshot(), state)
    np.testing.assert_array_equal(device_rng_as_pairs(rt.rng_states), rng)


def test_he3_cartesian_float64_trajectory_within_1e12(gpu):
    from models.he3.model import He3Model
    model = He3Model(pressure=20, g_d=0.001, dtype=np.complex128)
    rt = _runtime(model, 6, [0.3, 0.5], 228)
    state = _he3_random_state(6, 2, 6, np.complex128)
    rt.state.copy_to_device(state)
    rng = oracle_rng(rt)
    betas = (1.0 / rt.temperatures).astype(np.float64)
    for _ in range(100):
        rt.sweep()
        for color in range(4):
            oracle_color("he3", state, rng, betas, model, color)
    np.testing.assert_allclose(rt.snapshot(), state, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_he3_gaussian_trajectory_and_counts(gpu, dtype):
    from models.he3.model import He3Model
    from models.he3.sampling import He3GaussianRuntime
    model = He3Model(pressure=20, g_d=0.001, dtype=dtype)
    rt = He3GaussianRuntime(model, 6, [0.3, 0.4, 0.6], threads=2, seed=227)
    state = model.initialize_state(6, 3, 227).copy()
    rng = oracle_rng(rt)
    betas = (1.0 / rt.temperatures).astype(model.beta_dtype)
    counts = np.zeros((3, 3, 6, 6), dtype=np.int64)
    for _ in range(12):
        rt.sweep()
        for color in range(4):
            oracle.he3_gaussian_color(state, rng, betas, color, model.betas, model.K,
                                      rt.gaussian.rho, rt.gaussian.mass, counts, model.g_d)
    if dtype == np.complex64:
        np.testing.assert_array_equal(rt.snapshot(), state)
    else:
        np.testing.assert_allclose(rt.snapshot(), state, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(rt.gaussian.counts.copy_to_host(), counts)
    diag = rt.sampling_diagnostics()
    assert sum(diag["nonfinite_rejections"]) == 0
    assert sum(diag["accepted"]) > 0


def _he3_total_energy(state, K, betas, twist=(0.0, 0.0), g_d=0.0):
    psi = np.asarray(state, dtype=np.complex128)
    density = np.sum(np.abs(psi)**2, axis=(-2, -1))
    Q = np.einsum("...ma,...na->...mn", psi, psi)
    G = np.einsum("...ma,...na->...mn", psi, psi.conj())
    inv = (np.abs(np.trace(Q, axis1=-2, axis2=-1))**2, density**2,
           np.sum(np.abs(Q)**2, axis=(-2, -1)), np.sum(np.abs(G)**2, axis=(-2, -1)),
           np.sum((G**2).real, axis=(-2, -1)))
    pot = -2 * density + sum(b * x for b, x in zip(betas, inv))
    matrix = np.zeros(psi.shape[:-2] + (3, 3), dtype=np.complex128)
    matrix[..., :, :2] = psi
    dip = (np.abs(np.trace(matrix, axis1=-2, axis2=-1))**2
           + np.einsum("...ij,...ji->...", matrix, matrix.conj()).real
           - (2 / 3) * np.sum(np.abs(matrix)**2, axis=(-2, -1)))
    pot = pot + g_d * dip
    dx = np.exp(-1j * twist[0]) * np.roll(psi, -1, axis=1) - psi
    dy = np.exp(-1j * twist[1]) * np.roll(psi, -1, axis=2) - psi
    G1 = np.sum(np.abs(dx)**2 + np.abs(dy)**2, axis=(-2, -1))
    G2 = np.sum(np.abs(dx[..., 0] + dy[..., 1])**2, axis=-1)
    G3 = np.sum(np.abs(dx[..., 0])**2 + np.abs(dy[..., 1])**2
                + 2 * (dx[..., 1].conj() * dy[..., 0]).real, axis=-1)
    return np.sum(pot + K * (G1 + G2 + G3), axis=(1, 2))


def test_he3_measurements_match_oracle_and_numpy(gpu):
    from models.he3.model import He3Model
    K = np.float32(0.73)
    betas = np.asarray((-0.4, 1.1, 0.8, 1.3, -0.7), dtype=np.float32)
    for g_d in (0.0, 0.1):
        model = He3Model(K=K, betas=betas, g_d=g_d)
        for L in (4, 6, 20, 80):
            state = _he3_random_state(L)
            dev = gpu.to_device(state)
            measured = model.measure(dev)
            ref = oracle.he3_measure(state, model.betas, model.K, model.g_d)
            for c, name in enumerate(measured):
                np.testing.assert_allclose(measured[name], ref[c], rtol=1e-12, atol=1e-9,
                                           err_msg=name)
            expected = _he3_total_energy(state, float(K), betas.astype(float), g_d=float(model.g_d))
            budget = 8 * np.finfo(np.float32).eps
            np.testing.assert_allclose(measured["energy"] / L**2, expected / L**2, atol=budget, rtol=0)
            energy = gpu.device_array(3, np.float64)
            model.measure_energy(dev, energy)
            np.testing.assert_array_equal(energy.copy_to_host(), measured["energy"])
            cached = model.measure_with_energy(dev, energy)
            for name in measured:
                np.testing.assert_array_equal(cached[name], measured[name])
            order = np.array([2, 0, 1], dtype=np.int32)
            ordered = model.measure(dev, gpu.to_device(order))
            for name in measured:
                np.testing.assert_array_equal(ordered[name][order], measured[name])


def test_he3_helicity_derivatives_match_finite_differences(gpu):
    from models.he3.model import He3Model
    K = np.float32(0.73)
    betas = np.asarray((-0.4, 1.1, 0.8, 1.3, -0.7), dtype=np.float32)
    state = _he3_random_state(L=4)
    measured = He3Model(K=K, betas=betas).measure(gpu.to_device(state))
    h = 2e-4
    energy = _he3_total_energy(state, float(K), betas.astype(float))
    for axis, direction in enumerate(("x", "y")):
        twist = np.zeros(2)
        twist[axis] = h
        plus = _he3_total_energy(state, float(K), betas.astype(float), twist)
        minus = _he3_total_energy(state, float(K), betas.astype(float), -twist)
        np.testing.assert_allclose(measured[f"current_{direction}"], (plus - minus) / (2 * h),
                                   atol=2e-6, rtol=1e-7)
        np.testing.assert_allclose(measured[f"curvature_{direction}"],
                                   (plus - 2 * energy + minus) / h**2, atol=2e-5, rtol=1e-6)


def test_he3_spin_alignment_special_values(gpu):
    from models.he3.model import He3Model
    state = np.zeros((5, 4, 4, 3, 2), dtype=np.complex64)
    orbital = np.array([1, 1j], dtype=np.complex64) / np.sqrt(2)
    state[0, ..., 2, :] = orbital
    state[1, ..., 0, :] = orbital
    state[2] = np.ones((3, 1)) * orbital / np.sqrt(3)
    state[3, :2, :, 2, :] = np.sqrt(3) * orbital
    state[3, 2:, :, 0, :] = orbital
    measured = He3Model(g_d=0.001).measure(gpu.to_device(state))
    np.testing.assert_allclose(measured["spin_alignment"], [1, -0.5, 0, 0.625, 0], atol=1e-7, rtol=0)


def test_he3_ordered_state_closed_forms(gpu):
    from models.he3.model import He3Model
    K, betas = 1.3, (-0.5, 1.0, 1.0, 1.0, -1.0)
    model = He3Model(K=K, betas=betas)
    rt = _runtime(model, 4, [0.6, 0.9], 42)
    state = rt.snapshot()
    B = sum(betas[i] for i in (1, 3, 4))
    n = float(np.sum(np.abs(state[0, 0, 0].astype(np.complex128))**2))
    np.testing.assert_allclose(n, 1 / B, atol=1e-7)
    measured = rt.measure()
    expected = {"energy": -2 * n * 16 + B * n**2 * 16, "density": n * 16, "chirality": n * 16,
                "phase_real": n * 16, "phase_imag": 0, "current_x": 0, "current_y": 0,
                "curvature_x": 4 * float(np.float32(K)) * n * 16,
                "curvature_y": 4 * float(np.float32(K)) * n * 16, "spin_alignment": 1}
    for name, value in expected.items():
        np.testing.assert_allclose(measured[name], value, atol=1e-6)
