"""numpy wrappers around the C oracle (test infrastructure only).

Each wrapper mutates its numpy arguments in place exactly as the reference
kernel mutates the device arrays it is given, so a test can run the oracle and
the HIP path side by side on copies of the same inputs.
"""

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liboracle.so"
_lib = None

__all__ = [
    "load", "build", "splitmix64", "rng_init", "rng_draw_u64", "rng_draw_f32",
    "rng_draw_f64", "rng_jump", "set_threads",
    "ising_color", "ising_sweeps", "ising_measure", "ising_energy",
    "xy_color", "xy_gauss_color", "xy_sweeps", "xy_measure", "xy_energy",
    "kawamura_local", "kawamura_color", "kawamura_sweeps", "kawamura_measure",
    "kawamura_energy",
    "he3_potential", "he3_gradient_difference", "he3_local_energy", "he3_color",
    "he3_sweeps", "he3_gaussian_color", "he3_measure", "he3_energy",
    "pt_exchange", "fast_sincosf", "fast_expf", "fast_logf", "normal_q",
    "accept_weight", "xyfast_cache", "xyfast_color", "xyfast_gauss_color", "fast_gauss_draw",
    "xyfast_sweeps", "xyfast_observables",
]


def build(force=False):
    """Compile liboracle.so with the committed Makefile."""
    if force or not _LIB_PATH.exists():
        subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def load():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(_LIB_PATH))
        _declare(_lib)
    return _lib


P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
U64 = ctypes.c_uint64
F32 = ctypes.c_float
F64 = ctypes.c_double


def _declare(lib):
    sig = {
        "orc_splitmix64": (U64, [U64]),
        "orc_rng_next": (U64, [P]),
        "orc_rng_jump": (None, [P]),
        "orc_rng_init": (None, [P, I64, U64, U64]),
        "orc_rng_draw_u64": (None, [P, I64, I64, P]),
        "orc_rng_draw_f32": (None, [P, I64, I64, P]),
        "orc_rng_draw_f64": (None, [P, I64, I64, P]),
        "orc_set_threads": (None, [ctypes.c_int]),
        "orc_ising_color": (None, [P, P, P, I32, I32, I32, F32]),
        "orc_ising_sweeps": (None, [P, P, P, I32, I32, F32, I64]),
        "orc_ising_measure": (None, [P, I32, I32, F32, P, P, P]),
        "orc_ising_energy": (None, [P, I32, I32, F32, P]),
        "orc_xy_color_f32": (None, [P, P, P, I32, I32, I32, F32, F32]),
        "orc_xy_color_f64": (None, [P, P, P, I32, I32, I32, F64, F64]),
        "orc_xy_gauss_color_f32": (None, [P, P, P, I32, I32, I32, F32, P, P, I32, P]),
        "orc_xy_gauss_color_f64": (None, [P, P, P, I32, I32, I32, F64, P, P, I32, P]),
        "orc_xy_sweeps_f32": (None, [P, P, P, I32, I32, F32, F32, I64]),
        "orc_xy_sweeps_f64": (None, [P, P, P, I32, I32, F64, F64, I64]),
        "orc_xy_measure_f32": (None, [P, I32, I32, F64, P, P, P]),
        "orc_xy_measure_f64": (None, [P, I32, I32, F64, P, P, P]),
        "orc_xy_energy_f32": (None, [P, I32, I32, F64, P]),
        "orc_xy_energy_f64": (None, [P, I32, I32, F64, P]),
        "orc_kawamura_local_f32": (F32, [P, I32, I32, I32, I32, F32, F32, F32, F32, F32, F32]),
        "orc_kawamura_local_f64": (F64, [P, I32, I32, I32, I32, F64, F64, F64, F64, F64, F64]),
        "orc_kawamura_color_f32": (None, [P, P, P, I32, I32, I32, F32, F32]),
        "orc_kawamura_color_f64": (None, [P, P, P, I32, I32, I32, F64, F64]),
        "orc_kawamura_sweeps_f32": (None, [P, P, P, I32, I32, F32, F32, I64]),
        "orc_kawamura_sweeps_f64": (None, [P, P, P, I32, I32, F64, F64, I64]),
        "orc_kawamura_measure_f32": (None, [P, I32, I32, F64, P, P, P]),
        "orc_kawamura_measure_f64": (None, [P, I32, I32, F64, P, P, P]),
        "orc_kawamura_energy_f32": (None, [P, I32, I32, F64, P]),
        "orc_kawamura_energy_f64": (None, [P, I32, I32, F64, P]),
        "orc_he3_potential_f32": (F32, [P, P, F32]),
        "orc_he3_potential_f64": (F64, [P, P, F64]),
        "orc_he3_gradient_difference_f32": (F32, [P, I32, I32, I32, I32, P]),
        "orc_he3_gradient_difference_f64": (F64, [P, I32, I32, I32, I32, P]),
        "orc_he3_local_energy_f32": (F32, [P, I32, I32, I32, I32, P, P, F32, F32]),
        "orc_he3_color_f32": (None, [P, P, P, I32, I32, I32, P, F32, F32, F32]),
        "orc_he3_color_f64": (None, [P, P, P, I32, I32, I32, P, F64, F64, F64]),
        "orc_he3_sweeps_f32": (None, [P, P, P, I32, I32, P, F32, F32, F32, I64]),
        "orc_he3_gaussian_color_f32": (None, [P, P, P, I32, I32, I32, P, F32, F32, F32, P, F32]),
        "orc_he3_gaussian_color_f64": (None, [P, P, P, I32, I32, I32, P, F64, F64, F64, P, F64]),
        "orc_he3_measure_f32": (None, [P, I32, I32, P, F32, F32, P, P, P]),
        "orc_he3_measure_f64": (None, [P, I32, I32, P, F64, F64, P, P, P]),
        "orc_he3_energy_f32": (None, [P, I32, I32, P, F32, F32, P]),
        "orc_he3_energy_f64": (None, [P, I32, I32, P, F64, F64, P]),
        "orc_pt_exchange_f32": (None, [P, P, P, P, P, P, P, I32, I32, I64]),
        "orc_fast_sincosf": (None, [F32, P, P]),
        "orc_fast_expf": (F32, [F32]),
        "orc_fast_sincosf_n": (None, [P, I64, P, P]),
        "orc_fast_expf_n": (None, [P, I64, P]),
        "orc_fast_logf_n": (None, [P, I64, P]),
        "orc_accept_weight": (F32, [F32]),
        "orc_xyfast_init_cache": (None, [P, I64, P, P]),
        "orc_xyfast_color": (None, [P, P, P, P, P, I32, I32, I32, F32, F32]),
        "orc_xyfast_sweeps": (None, [P, P, P, P, P, I32, I32, F32, F32, I64]),
        "orc_xyfast_gauss_color": (None, [P, P, P, P, P, P, I32, I32, I32, F32, P]),
        "orc_fast_gauss_draw": (None, [ctypes.c_uint64, P, P]),
        "orc_normal_q_n": (None, [P, I64, P]),
        "orc_xyfast_observables": (None, [P, P, I32, I32, F32, P]),
        "orc_pt_exchange_f64": (None, [P, P, P, P, P, P, P, I32, I32, I64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def _ptr(a, dtype=None, writable=False):
    if a is None:
        return None
    if not isinstance(a, np.ndarray):
        raise TypeError("oracle arguments must be numpy arrays")
    if dtype is not None and a.dtype != np.dtype(dtype):
        raise TypeError(f"expected {np.dtype(dtype)}, got {a.dtype}")
    if not a.flags.c_contiguous:
        raise ValueError("oracle arrays must be C-contiguous")
    if writable and not a.flags.writeable:
        raise ValueError("array must be writable")
    return a.ctypes.data


def _rng(a):
    if a.dtype != np.uint64 or a.shape[-1:] != (2,):
        raise TypeError("RNG states must be uint64 with a trailing axis of 2 (s0, s1)")
    return _ptr(a, np.uint64, True)


def _real(state):
    if state.dtype in (np.float32, np.complex64):
        return "f32", np.float32
    if state.dtype in (np.float64, np.complex128):
        return "f64", np.float64
    raise TypeError(f"unsupported state dtype {state.dtype}")


def set_threads(n):
    load().orc_set_threads(int(n))


# ---------------------------------------------------------------- RNG

def splitmix64(seed):
    return int(load().orc_splitmix64(int(seed) & (2**64 - 1)))


def rng_init(n, seed, subsequence_start=0):
    """numba create_xoroshiro128p_states(n, seed) as an (n, 2) uint64 array."""
    states = np.zeros((n, 2), dtype=np.uint64)
    if n:
        load().orc_rng_init(states.ctypes.data, int(n), int(seed) & (2**64 - 1),
                            int(subsequence_start))
    return states


def rng_jump(state):
    s = np.ascontiguousarray(state, dtype=np.uint64).reshape(2).copy()
    load().orc_rng_jump(s.ctypes.data)
    return s


def rng_draw_u64(states, index, count):
    out = np.empty(count, dtype=np.uint64)
    load().orc_rng_draw_u64(_rng(states), int(index), int(count), out.ctypes.data)
    return out


def rng_draw_f32(states, index, count):
    out = np.empty(count, dtype=np.float32)
    load().orc_rng_draw_f32(_rng(states), int(index), int(count), out.ctypes.data)
    return out


def rng_draw_f64(states, index, count):
    out = np.empty(count, dtype=np.float64)
    load().orc_rng_draw_f64(_rng(states), int(index), int(count), out.ctypes.data)
    return out


# -------------------------------------------------------------- Ising

def ising_color(spins, rng, betas, color, J):
    R, L = spins.shape[0], spins.shape[1]
    load().orc_ising_color(_ptr(spins, np.int8, True), _rng(rng), _ptr(betas, np.float32),
                           R, L, int(color), float(J))


def ising_sweeps(spins, rng, betas, J, nsweeps):
    R, L = spins.shape[0], spins.shape[1]
    load().orc_ising_sweeps(_ptr(spins, np.int8, True), _rng(rng), _ptr(betas, np.float32),
                            R, L, float(J), int(nsweeps))


def ising_measure(spins, J, temperature_indices=None, cached_energies=None):
    R, L = spins.shape[0], spins.shape[1]
    out = np.zeros((2, R), dtype=np.float64)
    load().orc_ising_measure(_ptr(spins, np.int8), R, L, float(J),
                             _ptr(temperature_indices, np.int32),
                             _ptr(cached_energies, np.float64), out.ctypes.data)
    return out


def ising_energy(spins, J):
    R, L = spins.shape[0], spins.shape[1]
    out = np.zeros(R, dtype=np.float64)
    load().orc_ising_energy(_ptr(spins, np.int8), R, L, float(J), out.ctypes.data)
    return out


# ----------------------------------------------------------------- XY

def xy_color(theta, rng, betas, color, J, delta):
    sfx, real = _real(theta)
    R, L = theta.shape[0], theta.shape[1]
    getattr(load(), f"orc_xy_color_{sfx}")(_ptr(theta, real, True), _rng(rng), _ptr(betas, real),
                                           R, L, int(color), float(J), float(delta))


def xy_gauss_color(theta, rng, betas, color, J, table_betas, sigmas, accepted=None):
    """Gaussian-proposal colour update; `accepted` (int64, per temperature) is incremented."""
    sfx, real = _real(theta)
    R, L = theta.shape[0], theta.shape[1]
    table_betas = np.ascontiguousarray(table_betas, dtype=real)
    sigmas = np.ascontiguousarray(sigmas, dtype=real)
    acc = None if accepted is None else _ptr(accepted, np.int64, True)
    getattr(load(), f"orc_xy_gauss_color_{sfx}")(_ptr(theta, real, True), _rng(rng), _ptr(betas, real), R, L,
                                                 int(color), float(J), table_betas.ctypes.data,
                                                 sigmas.ctypes.data, len(sigmas), acc)


def xy_sweeps(theta, rng, betas, J, delta, nsweeps):
    sfx, real = _real(theta)
    R, L = theta.shape[0], theta.shape[1]
    getattr(load(), f"orc_xy_sweeps_{sfx}")(_ptr(theta, real, True), _rng(rng), _ptr(betas, real),
                                            R, L, float(J), float(delta), int(nsweeps))


def xy_measure(theta, J, temperature_indices=None, cached_energies=None):
    sfx, real = _real(theta)
    R, L = theta.shape[0], theta.shape[1]
    out = np.zeros((5, R), dtype=np.float64)
    getattr(load(), f"orc_xy_measure_{sfx}")(_ptr(theta, real), R, L, float(J),
                                             _ptr(temperature_indices, np.int32),
                                             _ptr(cached_energies, np.float64), out.ctypes.data)
    return out


def xy_energy(theta, J):
    sfx, real = _real(theta)
    R, L = theta.shape[0], theta.shape[1]
    out = np.zeros(R, dtype=np.float64)
    getattr(load(), f"orc_xy_energy_{sfx}")(_ptr(theta, real), R, L, float(J), out.ctypes.data)
    return out


# ----------------------------------------------------------- Kawamura

def kawamura_local(state, r, i, j, site, J):
    sfx, real = _real(state)
    L = state.shape[1]
    return getattr(load(), f"orc_kawamura_local_{sfx}")(
        _ptr(state, real), L, int(r), int(i), int(j), *[float(v) for v in site], float(J))


def kawamura_color(state, rng, betas, color, J, delta):
    sfx, real = _real(state)
    R, L = state.shape[0], state.shape[1]
    getattr(load(), f"orc_kawamura_color_{sfx}")(_ptr(state, real, True), _rng(rng),
                                                 _ptr(betas, real), R, L, int(color),
                                                 float(J), float(delta))


def kawamura_sweeps(state, rng, betas, J, delta, nsweeps):
    sfx, real = _real(state)
    R, L = state.shape[0], state.shape[1]
    getattr(load(), f"orc_kawamura_sweeps_{sfx}")(_ptr(state, real, True), _rng(rng),
                                                  _ptr(betas, real), R, L, float(J),
                                                  float(delta), int(nsweeps))


def kawamura_measure(state, J, temperature_indices=None, cached_energies=None):
    sfx, real = _real(state)
    R, L = state.shape[0], state.shape[1]
    out = np.zeros((7, R), dtype=np.float64)
    getattr(load(), f"orc_kawamura_measure_{sfx}")(_ptr(state, real), R, L, float(J),
                                                   _ptr(temperature_indices, np.int32),
                                                   _ptr(cached_energies, np.float64),
                                                   out.ctypes.data)
    return out


def kawamura_energy(state, J):
    sfx, real = _real(state)
    R, L = state.shape[0], state.shape[1]
    out = np.zeros(R, dtype=np.float64)
    getattr(load(), f"orc_kawamura_energy_{sfx}")(_ptr(state, real), R, L, float(J),
                                                  out.ctypes.data)
    return out


# ---------------------------------------------------------------- He3

def _coef(coef, real):
    return np.ascontiguousarray(np.asarray(coef, dtype=real).reshape(5))


def he3_potential(psi, coef, g_d):
    sfx, real = _real(psi)
    psi = np.ascontiguousarray(psi).reshape(3, 2)
    c = _coef(coef, real)
    return getattr(load(), f"orc_he3_potential_{sfx}")(psi.ctypes.data, c.ctypes.data, float(g_d))


def he3_gradient_difference(state, r, i, j, proposed):
    sfx, real = _real(state)
    prop = np.ascontiguousarray(proposed, dtype=state.dtype).reshape(3, 2)
    return getattr(load(), f"orc_he3_gradient_difference_{sfx}")(
        _ptr(state, state.dtype), state.shape[1], int(r), int(i), int(j), prop.ctypes.data)


def he3_local_energy(state, r, i, j, psi, coef, K, g_d):
    psi = np.ascontiguousarray(psi, dtype=np.complex64).reshape(3, 2)
    c = _coef(coef, np.float32)
    return load().orc_he3_local_energy_f32(_ptr(state, np.complex64), state.shape[1], int(r),
                                           int(i), int(j), psi.ctypes.data, c.ctypes.data,
                                           float(K), float(g_d))


def he3_color(state, rng, betas, color, coef, K, delta, g_d):
    sfx, real = _real(state)
    R, L = state.shape[0], state.shape[1]
    c = _coef(coef, real)
    getattr(load(), f"orc_he3_color_{sfx}")(_ptr(state, state.dtype, True), _rng(rng),
                                            _ptr(betas, real), R, L, int(color), c.ctypes.data,
                                            float(K), float(delta), float(g_d))


def he3_sweeps(state, rng, betas, coef, K, delta, g_d, nsweeps):
    R, L = state.shape[0], state.shape[1]
    c = _coef(coef, np.float32)
    load().orc_he3_sweeps_f32(_ptr(state, np.complex64, True), _rng(rng),
                              _ptr(betas, np.float32), R, L, c.ctypes.data, float(K),
                              float(delta), float(g_d), int(nsweeps))


def he3_gaussian_color(state, rng, betas, color, coef, K, rho, mass, counts, g_d):
    sfx, real = _real(state)
    R, L = state.shape[0], state.shape[1]
    c = _coef(coef, real)
    getattr(load(), f"orc_he3_gaussian_color_{sfx}")(
        _ptr(state, state.dtype, True), _rng(rng), _ptr(betas, real), R, L, int(color),
        c.ctypes.data, float(K), float(rho), float(mass), _ptr(counts, np.int64, True),
        float(g_d))


def he3_measure(state, coef, K, g_d, temperature_indices=None, cached_energies=None):
    sfx, real = _real(state)
    R, L = state.shape[0], state.shape[1]
    c = _coef(coef, real)
    out = np.zeros((10, R), dtype=np.float64)
    getattr(load(), f"orc_he3_measure_{sfx}")(_ptr(state, state.dtype), R, L, c.ctypes.data,
                                              float(K), float(g_d),
                                              _ptr(temperature_indices, np.int32),
                                              _ptr(cached_energies, np.float64), out.ctypes.data)
    return out


def he3_energy(state, coef, K, g_d):
    sfx, real = _real(state)
    R, L = state.shape[0], state.shape[1]
    c = _coef(coef, real)
    out = np.zeros(R, dtype=np.float64)
    getattr(load(), f"orc_he3_energy_{sfx}")(_ptr(state, state.dtype), R, L, c.ctypes.data,
                                             float(K), float(g_d), out.ctypes.data)
    return out


# ----------------------------------------------------------------- PT

def pt_exchange(energies, betas, w2t, t2w, pt_rng, counts, trips, parity, sweep):
    R = energies.shape[0]
    sfx = "f32" if betas.dtype == np.float32 else "f64"
    getattr(load(), f"orc_pt_exchange_{sfx}")(
        _ptr(energies, np.float64), _ptr(betas, betas.dtype, True), _ptr(w2t, np.int32, True),
        _ptr(t2w, np.int32, True), _rng(pt_rng), _ptr(counts, np.int64, True),
        _ptr(trips, np.int64, True), R, int(parity), int(sweep))


# ------------------------------------------------ batched float32 fast path

def fast_sincosf(x):
    """(sin, cos) arrays from the oracle's float32 routine."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    sn = np.empty_like(x)
    cs = np.empty_like(x)
    load().orc_fast_sincosf_n(x.ctypes.data, x.size, sn.ctypes.data, cs.ctypes.data)
    return sn, cs


def fast_expf(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    load().orc_fast_expf_n(x.ctypes.data, x.size, y.ctypes.data)
    return y


def accept_weight(x):
    f = load().orc_accept_weight
    x = np.ascontiguousarray(x, dtype=np.float32)
    return np.array([f(float(v)) for v in x.reshape(-1)], dtype=np.float32).reshape(x.shape)


def fast_logf(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.empty_like(x)
    load().orc_fast_logf_n(x.ctypes.data, x.size, y.ctypes.data)
    return y


def normal_q(k):
    """Q((2k + 1) 2^-23) = -Phi^-1 of the Gaussian step for 21-bit integers k."""
    k = np.ascontiguousarray(k, dtype=np.uint32)
    y = np.empty(k.shape, np.float32)
    load().orc_normal_q_n(k.ctypes.data, k.size, y.ctypes.data)
    return y


def xyfast_cache(theta):
    """(cos, sin) cache arrays of the fast XY update."""
    c = np.empty_like(theta, dtype=np.float32)
    s = np.empty_like(theta, dtype=np.float32)
    load().orc_xyfast_init_cache(_ptr(theta, np.float32), theta.size, c.ctypes.data, s.ctypes.data)
    return c, s


def xyfast_color(theta, c, s, rng, betas, color, J, delta):
    R, L = theta.shape[0], theta.shape[1]
    load().orc_xyfast_color(_ptr(theta, np.float32, True), _ptr(c, np.float32, True),
                            _ptr(s, np.float32, True), _rng(rng), _ptr(betas, np.float32), R, L,
                            int(color), float(J), float(delta))


def xyfast_gauss_color(theta, c, s, rng, betas, sigmas, color, J, accepted=None):
    """Fast-path Gaussian colour update; sigmas and accepted are per walker."""
    R, L = theta.shape[0], theta.shape[1]
    acc = None if accepted is None else _ptr(accepted, np.int64, True)
    sig = np.ascontiguousarray(sigmas, np.float32)  # keep alive across the call
    load().orc_xyfast_gauss_color(_ptr(theta, np.float32, True), _ptr(c, np.float32, True),
                                  _ptr(s, np.float32, True), _rng(rng), _ptr(betas, np.float32),
                                  _ptr(sig, np.float32), R, L, int(color), float(J), acc)


def fast_gauss_draw(x):
    """(z, v) of the batched engine's Gaussian step for 64-bit draws x."""
    x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1)
    z = np.empty(len(x), np.float32)
    v = np.empty(len(x), np.float32)
    zf, vf = F32(), F32()
    f = load().orc_fast_gauss_draw
    for k, xv in enumerate(x):
        f(int(xv), ctypes.byref(zf), ctypes.byref(vf))
        z[k], v[k] = zf.value, vf.value
    return z, v


def xyfast_sweeps(theta, c, s, rng, betas, J, delta, nsweeps):
    R, L = theta.shape[0], theta.shape[1]
    load().orc_xyfast_sweeps(_ptr(theta, np.float32, True), _ptr(c, np.float32, True),
                             _ptr(s, np.float32, True), _rng(rng), _ptr(betas, np.float32), R, L,
                             float(J), float(delta), int(nsweeps))


def xyfast_observables(c, s, J):
    R, L = c.shape[0], c.shape[1]
    out = np.zeros((5, R), dtype=np.float64)
    load().orc_xyfast_observables(_ptr(c, np.float32), _ptr(s, np.float32), R, L, float(J),
                                  out.ctypes.data)
    return out
