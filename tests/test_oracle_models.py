"""Pins of the CPU oracle's model functions against the reference's own test
expectations (closed forms, independent NumPy float64 Hamiltonians, exact
limits) -- the same checks the reference runs on its GPU kernels
(tests/test_simulation.py, tests/test_kawamura.py, tests/test_he3.py,
tests/test_he3_energy_difference.py, tests/test_parallel_tempering.py)."""

import numpy as np
import pytest

import oracle


# -------------------------------------------------------------------- Ising

def test_ising_measure_closed_forms_and_bond_count():
    J = np.float32(1.37)
    L = 6
    checkerboard = 1 - 2 * (np.indices((L, L)).sum(axis=0) % 2)
    spins = np.stack((np.ones((L, L)), -np.ones((L, L)), checkerboard)).astype(np.int8)
    out = oracle.ising_measure(spins, J)
    np.testing.assert_array_equal(out[0], float(J) * L**2 * np.array([-2.0, -2.0, 2.0]))
    np.testing.assert_array_equal(out[1], [L**2, -L**2, 0])
    rng = np.random.default_rng(18)
    for L in (4, 6, 32, 64):
        spins = rng.choice(np.array([-1, 1], dtype=np.int8), (3, L, L))
        equal = (np.sum(spins == np.roll(spins, -1, axis=1), axis=(1, 2))
                 + np.sum(spins == np.roll(spins, -1, axis=2), axis=(1, 2)))
        out = oracle.ising_measure(spins, J)
        np.testing.assert_array_equal(out[0], float(J) * (2 * L**2 - 2 * equal))
        np.testing.assert_array_equal(out[1], spins.sum(axis=(1, 2)))
        np.testing.assert_array_equal(oracle.ising_energy(spins, J), out[0])


def test_ising_temperature_limits_are_exact():
    L = 6
    rng = np.random.default_rng(19)
    spins = rng.choice(np.array([-1, 1], dtype=np.int8), (2, L, L))
    expected = spins.copy()
    states = oracle.rng_init(2 * L * L, 42)
    betas = np.array([1e8, 0.0], dtype=np.float32)
    for color in range(4):
        for r in range(2):
            for i in range(color // 2, L, 2):
                for j in range(color % 2, L, 2):
                    nb = (int(expected[r, (i - 1) % L, j]) + int(expected[r, (i + 1) % L, j])
                          + int(expected[r, i, (j - 1) % L]) + int(expected[r, i, (j + 1) % L]))
                    if r == 1 or expected[r, i, j] * nb <= 0:
                        expected[r, i, j] *= -1
        oracle.ising_color(spins, states, betas, color, np.float32(1.37))
        np.testing.assert_array_equal(spins, expected)


def test_ising_draws_one_uniform_per_site_visit():
    L, R = 4, 2
    states = oracle.rng_init(R * L * L, 5)
    before = states.copy()
    spins = np.ones((R, L, L), dtype=np.int8)
    oracle.ising_sweeps(spins, states, np.array([0.5, 0.7], np.float32), 1.0, 3)
    for k in range(R * L * L):
        s = before[k:k + 1].copy()
        oracle.rng_draw_u64(s, 0, 3)
        np.testing.assert_array_equal(states[k], s[0])


# ----------------------------------------------------------------------- XY

def _xy_reference(angles, J):
    spins = np.exp(1j * angles.astype(np.float64))
    bx = spins * np.conj(np.roll(spins, -1, axis=1))
    by = spins * np.conj(np.roll(spins, -1, axis=2))
    return np.stack([-J * np.sum(bx.real + by.real, axis=(1, 2)), spins.real.sum(axis=(1, 2)),
                     spins.imag.sum(axis=(1, 2)), bx.real.sum(axis=(1, 2)), bx.imag.sum(axis=(1, 2))])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_xy_measure_matches_complex_exponential_reference(dtype):
    rng = np.random.default_rng(17)
    for L in (4, 6, 32):
        angles = rng.uniform(-np.pi, np.pi, (3, L, L)).astype(dtype)
        np.testing.assert_allclose(oracle.xy_measure(angles, 1.37), _xy_reference(angles, 1.37),
                                   rtol=1e-12, atol=1e-10)
        np.testing.assert_allclose(oracle.xy_energy(angles, 1.37), _xy_reference(angles, 1.37)[0],
                                   rtol=1e-12, atol=1e-10)
    out = oracle.xy_measure(np.zeros((2, 4, 4), dtype), 1.37)
    np.testing.assert_allclose(out[:, 0], [-2 * 1.37 * 16, 16, 0, 16, 0], atol=1e-12)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_xy_updates_keep_angles_wrapped_and_obey_limits(dtype):
    L, R = 8, 2
    rng = np.random.default_rng(4)
    theta = rng.uniform(-np.pi, np.pi, (R, L, L)).astype(dtype)
    states = oracle.rng_init(R * L * L, 3)
    oracle.xy_sweeps(theta, states, np.zeros(R, dtype), 1.0, np.pi, 30)
    pi = dtype(np.pi)
    assert np.all(theta >= -pi) and np.all(theta < pi)
    assert np.std(theta) > 1.0
    cold = np.zeros((R, L, L), dtype)
    oracle.xy_sweeps(cold, states, np.full(R, 1e8, dtype), 1.0, 1.5, 5)
    np.testing.assert_array_equal(cold, 0)


# ----------------------------------------------------------------- Kawamura

def _kawamura_reference(state, J):
    theta, tau, d = state[..., 0], state[..., 1], state[..., 2:]
    phase = np.exp(1j * theta)
    bonds = []
    for axis in (1, 2):
        weight = (1 + tau * np.roll(tau, -1, axis)) * np.sum(d * np.roll(d, -1, axis), axis=-1)
        bonds.append(weight * phase * np.conj(np.roll(phase, -1, axis)))
    tensor = np.sum(d[..., :, None] * d[..., None, :] - np.eye(3) / 3, axis=(1, 2))
    return np.stack([-J * np.sum(bonds[0].real + bonds[1].real, axis=(1, 2)), tau.sum(axis=(1, 2)),
                     (phase**2).real.sum(axis=(1, 2)), (phase**2).imag.sum(axis=(1, 2)),
                     np.sum(tensor**2, axis=(1, 2)) / 3, bonds[0].real.sum(axis=(1, 2)),
                     bonds[0].imag.sum(axis=(1, 2))])


def _kawamura_state(L=6, R=3, seed=63, dtype=np.float32):
    rng = np.random.default_rng(seed)
    state = np.empty((R, L, L, 5), dtype=dtype)
    state[..., 0] = rng.uniform(-np.pi, np.pi, state.shape[:3])
    state[..., 1] = rng.choice([-1, 1], state.shape[:3])
    d = rng.normal(size=(*state.shape[:3], 3))
    state[..., 2:] = d / np.linalg.norm(d, axis=-1, keepdims=True)
    return state


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_kawamura_measure_energy_and_ordered_forms(dtype):
    J = 1.37
    for L in (4, 6, 20):
        state = _kawamura_state(L, dtype=dtype)
        expected = _kawamura_reference(state.astype(float), J)
        np.testing.assert_allclose(oracle.kawamura_measure(state, J), expected, atol=1e-10, rtol=1e-12)
        np.testing.assert_allclose(oracle.kawamura_energy(state, J), expected[0], atol=1e-10, rtol=1e-12)
    ordered = np.zeros((2, 4, 4, 5), dtype)
    ordered[..., 1] = 1
    ordered[..., 4] = 1
    np.testing.assert_allclose(oracle.kawamura_measure(ordered, J)[:, 0],
                               [-4 * J * 16, 16, 16, 0, 2 * 16**2 / 9, 32, 0], atol=1e-12)


def test_kawamura_local_energy_changes_match_total_hamiltonian():
    J = np.float32(1.37)
    state = _kawamura_state()
    old = _kawamura_reference(state.astype(float), float(J))[0]
    for field in ("phase", "chirality", "spin"):
        proposed = state[:, 1, 2].copy()
        if field == "phase":
            proposed[:, 0] += 0.7
        elif field == "chirality":
            proposed[:, 1] *= -1
        else:
            proposed[:, 2:] = np.array([0.6, 0, 0.8])
        changed = state.copy()
        changed[:, 1, 2] = proposed
        diff = _kawamura_reference(changed.astype(float), float(J))[0] - old
        for r in range(3):
            e0 = oracle.kawamura_local(state, r, 1, 2, state[r, 1, 2], J)
            e1 = oracle.kawamura_local(state, r, 1, 2, proposed[r], J)
            np.testing.assert_allclose(e1 - e0, diff[r], atol=3e-6, rtol=1e-5)


def test_kawamura_gauge_symmetry_and_update_constraints():
    state = _kawamura_state()
    original = oracle.kawamura_measure(state, 1.0)
    state[:, ::2, ::2, 0] += np.float32(np.pi)
    state[:, ::2, ::2, 2:] *= -1
    np.testing.assert_allclose(oracle.kawamura_measure(state, 1.0), original, atol=1e-5)
    hot = _kawamura_state(6, 2, seed=3)
    states = oracle.rng_init(2 * 36, 1)
    oracle.kawamura_sweeps(hot, states, np.zeros(2, np.float32), 1.0, 1.5, 20)
    assert set(np.unique(hot[..., 1])) == {-1.0, 1.0}
    np.testing.assert_allclose(np.linalg.norm(hot[..., 2:], axis=-1), 1, atol=2e-6)
    cold = np.zeros((2, 4, 4, 5), np.float32)
    cold[..., 1] = 1
    cold[..., 4] = 1
    oracle.kawamura_sweeps(cold, oracle.rng_init(32, 2), np.full(2, 1e8, np.float32), 1.0, 1.5, 4)
    np.testing.assert_allclose(oracle.kawamura_measure(cold, 1.0)[0], -64, atol=1e-5)


# ---------------------------------------------------------------------- He3

def _he3_state(L=6, R=3, seed=253, dtype=np.complex64, scale=0.3):
    rng = np.random.default_rng(seed)
    shape = (R, L, L, 3, 2)
    return (scale * (rng.normal(size=shape) + 1j * rng.normal(size=shape))).astype(dtype)


def he3_total_energy(state, K, betas, twist=(0.0, 0.0), g_d=0.0):
    """The reference test's independent float64 Hamiltonian (spin-matrix invariants)."""
    psi = np.asarray(state, dtype=np.complex128)
    density = np.sum(np.abs(psi)**2, axis=(-2, -1))
    Q = np.einsum("...ma,...na->...mn", psi, psi)
    G = np.einsum("...ma,...na->...mn", psi, psi.conj())
    inv = (np.abs(np.trace(Q, axis1=-2, axis2=-1))**2, density**2, np.sum(np.abs(Q)**2, axis=(-2, -1)),
           np.sum(np.abs(G)**2, axis=(-2, -1)), np.sum((G**2).real, axis=(-2, -1)))
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
    G3 = np.sum(np.abs(dx[..., 0])**2 + np.abs(dy[..., 1])**2 + 2 * (dx[..., 1].conj() * dy[..., 0]).real,
                axis=-1)
    return np.sum(pot + K * (G1 + G2 + G3), axis=(1, 2))


def test_he3_total_energy_and_measurements_against_spin_matrix_invariants():
    K = np.float32(0.73)
    betas = np.asarray((-0.4, 1.1, 0.8, 1.3, -0.7), dtype=np.float32)
    for g_d in (0.0, 0.1):
        for L in (4, 6, 20):
            state = _he3_state(L)
            out = oracle.he3_measure(state, betas, K, np.float32(g_d))
            expected = he3_total_energy(state, float(K), betas.astype(float), g_d=float(np.float32(g_d)))
            np.testing.assert_allclose(out[0] / L**2, expected / L**2, atol=8 * np.finfo(np.float32).eps, rtol=0)
            np.testing.assert_array_equal(oracle.he3_energy(state, betas, K, np.float32(g_d)), out[0])
            psi = state.astype(np.complex128)
            phase = np.sum(psi[..., 0]**2 - psi[..., 1]**2, axis=-1)
            eps2 = 2 * np.finfo(np.float32).eps
            np.testing.assert_allclose(out[1] / L**2, np.sum(np.abs(psi)**2, axis=(1, 2, 3, 4)) / L**2, atol=eps2)
            np.testing.assert_allclose(out[2] / L**2, 2 * np.sum((psi[..., 0].conj() * psi[..., 1]).imag,
                                                                 axis=(1, 2, 3)) / L**2, atol=eps2)
            np.testing.assert_allclose(out[3] / L**2, phase.real.sum(axis=(1, 2)) / L**2, atol=eps2)
            np.testing.assert_allclose(out[4] / L**2, phase.imag.sum(axis=(1, 2)) / L**2, atol=eps2)
    # float64 mode against the same Hamiltonian, tight
    state = _he3_state(6, dtype=np.complex128)
    out = oracle.he3_measure(state, betas.astype(float), 0.73, 0.1)
    np.testing.assert_allclose(out[0], he3_total_energy(state, 0.73, betas.astype(float), g_d=0.1),
                               rtol=1e-12, atol=1e-10)


def test_he3_twist_derivatives_match_central_differences():
    K = np.float32(0.73)
    betas = np.asarray((-0.4, 1.1, 0.8, 1.3, -0.7), dtype=np.float32)
    state = _he3_state(L=4)
    out = oracle.he3_measure(state, betas, K, 0.0)
    h = 2e-4
    e0 = he3_total_energy(state, float(K), betas.astype(float))
    for axis in range(2):
        tw = np.zeros(2)
        tw[axis] = h
        plus = he3_total_energy(state, float(K), betas.astype(float), tw)
        minus = he3_total_energy(state, float(K), betas.astype(float), -tw)
        np.testing.assert_allclose(out[5 + 2 * axis], (plus - minus) / (2 * h), atol=2e-6, rtol=1e-7)
        np.testing.assert_allclose(out[6 + 2 * axis], (plus - 2 * e0 + minus) / h**2, atol=2e-5, rtol=1e-6)


def test_he3_spin_alignment_special_values():
    state = np.zeros((5, 4, 4, 3, 2), dtype=np.complex64)
    orbital = np.array([1, 1j], dtype=np.complex64) / np.sqrt(2)
    state[0, ..., 2, :] = orbital
    state[1, ..., 0, :] = orbital
    state[2] = np.ones((3, 1)) * orbital / np.sqrt(3)
    state[3, :2, :, 2, :] = np.sqrt(3) * orbital
    state[3, 2:, :, 0, :] = orbital
    out = oracle.he3_measure(state, (-0.5, 1, 1, 1, -1), 1.0, 0.001)
    np.testing.assert_allclose(out[9], [1, -0.5, 0, 0.625, 0], atol=1e-7, rtol=0)


def _energy_parts(state, betas):
    psi = np.asarray(state, dtype=np.complex128)
    n = np.sum(np.abs(psi)**2, axis=(-2, -1))
    S = np.einsum("...ma,...na->...mn", psi, psi)
    H = np.einsum("...ma,...na->...mn", psi, psi.conj())
    inv = (np.abs(np.trace(S, axis1=-2, axis2=-1))**2, n**2, np.sum(np.abs(S)**2, axis=(-2, -1)),
           np.sum(np.abs(H)**2, axis=(-2, -1)), np.sum((H**2).real, axis=(-2, -1)))
    onsite = -2 * n + sum(float(b) * x for b, x in zip(betas, inv))
    dx = np.roll(psi, -1, axis=1) - psi
    dy = np.roll(psi, -1, axis=2) - psi
    g = (np.sum(np.abs(dx)**2 + np.abs(dy)**2, axis=(-2, -1))
         + np.sum(np.abs(dx[..., 0] + dy[..., 1])**2, axis=-1)
         + np.sum(np.abs(dx[..., 0])**2 + np.abs(dy[..., 1])**2 + 2 * (dx[..., 1].conj() * dy[..., 0]).real,
                  axis=-1))
    return onsite.sum(axis=(1, 2)), g.sum(axis=(1, 2))


def test_he3_gradient_difference_is_exact_change_of_the_three_cells():
    rng = np.random.default_rng(20261027)
    for L in (4, 6, 20):
        positions = [(0, 0), (0, L - 1), (L - 1, 0), (L - 1, L - 1), (L // 2, 0), (1, 2)]
        state = (0.3 * (rng.normal(size=(len(positions), L, L, 3, 2))
                        + 1j * rng.normal(size=(len(positions), L, L, 3, 2)))).astype(np.complex64)
        changed = state.copy()
        measured = []
        for r, (i, j) in enumerate(positions):
            prop = state[r, i, j].copy()
            prop.real += rng.uniform(-0.15, 0.15, size=(3, 2)).astype(np.float32)
            prop.imag += rng.uniform(-0.15, 0.15, size=(3, 2)).astype(np.float32)
            changed[r, i, j] = prop
            measured.append(oracle.he3_gradient_difference(state, r, i, j, prop))
        betas = (-0.5, 1, 1, 1, -1)
        _, g0 = _energy_parts(state, betas)
        _, g1 = _energy_parts(changed, betas)
        np.testing.assert_allclose(measured, g1 - g0, rtol=3e-6, atol=6e-6)
    # homogeneous field: the change is 8|e|^2 + 4 Re(conj(ex) ey) per spin row
    old = (0.7 * (rng.normal(size=(3, 2)) + 1j * rng.normal(size=(3, 2)))).astype(np.complex64)
    state = np.broadcast_to(old, (1, 6, 6, 3, 2)).copy()
    prop = old.copy()
    prop.real += np.float32(2**-18)
    step = prop.astype(np.complex128) - old.astype(np.complex128)
    expected = 8 * np.sum(np.abs(step)**2) + 4 * np.sum((step[:, 0].conj() * step[:, 1]).real)
    np.testing.assert_allclose(oracle.he3_gradient_difference(state, 0, 2, 3, prop), expected,
                               rtol=3e-6, atol=1e-13)


def test_he3_cartesian_limits():
    from models.he3.model import He3Model
    model = He3Model(delta=0.15)
    state = model.initialize_state(4, 2, 0)
    before = oracle.he3_measure(state, model.betas, model.K, model.g_d)[0]
    oracle.he3_sweeps(state, oracle.rng_init(32, 1), np.full(2, 1e8, np.float32), model.betas, model.K,
                      model.delta, model.g_d, 4)
    np.testing.assert_allclose(oracle.he3_measure(state, model.betas, model.K, model.g_d)[0], before, atol=1e-5)
    hot = model.initialize_state(6, 2, 0)
    init = hot.copy()
    oracle.he3_sweeps(hot, oracle.rng_init(72, 2), np.zeros(2, np.float32), model.betas, model.K,
                      model.delta, model.g_d, 10)
    assert np.isfinite(hot).all() and np.std(hot - init) > 0.03


# ----------------------------------------------------------------------- PT

def _pt_arrays(R):
    trips = np.zeros((5, R), dtype=np.int64)
    trips[0] = -1
    trips[2] = -1
    trips[0, 0], trips[0, -1] = 0, R - 1
    trips[2, [0, -1]] = 0
    return (np.arange(R, dtype=np.int32), np.arange(R, dtype=np.int32),
            np.zeros((3, R - 1), dtype=np.int64), trips)


def test_pt_equal_energies_always_swap_and_alternate_parity():
    for R in (4, 5, 259):
        betas = (1 / np.linspace(0.6, 1.4, R)).astype(np.float32)
        ladder = betas.copy()
        w2t, t2w, counts, trips = _pt_arrays(R)
        rng = oracle.rng_init(R - 1, 97)
        order = np.arange(R)
        energies = np.full(R, -3.0)
        for rnd in range(6):
            parity = rnd % 2 if R > 2 else 0
            for t in range(parity, R - 1, 2):
                order[t:t + 2] = order[t:t + 2][::-1]
            oracle.pt_exchange(energies, betas, w2t, t2w, rng, counts, trips, parity, rnd)
            np.testing.assert_array_equal(t2w, order)
            np.testing.assert_array_equal(w2t[t2w], np.arange(R))
            np.testing.assert_array_equal(betas, ladder[w2t])
        np.testing.assert_array_equal(counts[0], counts[1])


def test_pt_round_trip_bookkeeping_two_and_three_walkers():
    # two walkers, exchanges after sweeps 2, 4, 6 (tests/test_parallel_tempering.py:211-234)
    betas = np.array([1.0, 0.5], np.float32)
    w2t, t2w, counts, trips = _pt_arrays(2)
    rng = oracle.rng_init(1, 97)
    for sweep in (2, 4, 6):
        oracle.pt_exchange(np.zeros(2), betas, w2t, t2w, rng, counts, trips, 0, sweep)
    np.testing.assert_array_equal(trips[3], [1, 1])
    np.testing.assert_array_equal(trips[4], [4, 4])
    # three walkers, exchange every sweep (tests/test_parallel_tempering.py:236-253)
    betas = (1 / np.array([0.6, 1.0, 1.4])).astype(np.float32)
    w2t, t2w, counts, trips = _pt_arrays(3)
    rng = oracle.rng_init(2, 97)
    for sweep in range(1, 7):
        oracle.pt_exchange(np.zeros(3), betas, w2t, t2w, rng, counts, trips, (sweep - 1) % 2, sweep)
    np.testing.assert_array_equal(trips[3], [1, 0, 1])
    oracle.pt_exchange(np.zeros(3), betas, w2t, t2w, rng, counts, trips, 0, 7)
    np.testing.assert_array_equal(trips[3], [1, 1, 1])
    np.testing.assert_array_equal(trips[4], [5, 6, 6])


def test_pt_nonfinite_and_unfavourable_swaps():
    betas = np.array([1e6, 1e-6], np.float32)
    w2t, t2w, counts, trips = _pt_arrays(2)
    rng = oracle.rng_init(1, 98)
    before = rng.copy()
    oracle.pt_exchange(np.array([-43.84, 43.84]), betas, w2t, t2w, rng, counts, trips, 0, 1)
    np.testing.assert_array_equal(t2w, [0, 1])
    np.testing.assert_array_equal(counts[:, 0], [1, 0, 0])
    assert not np.array_equal(rng, before)  # a finite pair always draws
    before = rng.copy()
    oracle.pt_exchange(np.array([np.nan, 1.0]), betas, w2t, t2w, rng, counts, trips, 0, 2)
    np.testing.assert_array_equal(counts[:, 0], [2, 0, 1])
    np.testing.assert_array_equal(rng, before)  # a nonfinite pair draws nothing
    oracle.pt_exchange(np.array([43.84, -43.84]), betas, w2t, t2w, rng, counts, trips, 0, 3)
    np.testing.assert_array_equal(t2w, [1, 0])
