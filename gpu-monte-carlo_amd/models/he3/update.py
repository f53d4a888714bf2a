"""He3 Cartesian Metropolis colour update (models/he3/update.py:176-227 of the reference)."""

import ctypes

import numpy as np

from mcgpu import _lib, as_pointer
from models._device import STREAM, check_update_args


# He3Model keeps its coefficients in a tuple, which cannot change, so the
# float64 copy handed to the kernels is made once per tuple, not per launch.
_COEF_CACHE = {}


def coefficients(betas):
    """Host pointer to the five GL coefficients as float64 (kept alive by the cache or the pointer)."""
    if isinstance(betas, tuple):
        hit = _COEF_CACHE.get(betas)
        if hit is not None:
            return hit[1]
    arr = np.ascontiguousarray(np.asarray(betas, dtype=np.float64).reshape(5))
    ptr = arr.ctypes.data_as(ctypes.c_void_p)
    if isinstance(betas, tuple) and len(_COEF_CACHE) < 256:
        _COEF_CACHE[betas] = (arr, ptr)
    return ptr


def check_he3_state(state):
    if state.ndim != 5 or state.shape[3:] != (3, 2):
        raise ValueError("He3 state must have shape (R, L, L, 3, 2)")


def metropolis_color(state, rng_states, inverse_temperatures, color, betas, K, delta, g_d):
    """Six complex fields, each moved by a uniform box of half-width delta."""
    R, L, code = check_update_args(state, rng_states, inverse_temperatures, color)
    check_he3_state(state)
    _lib.call("mc_he3_update_color", STREAM, code, as_pointer(state), as_pointer(rng_states),
              as_pointer(inverse_temperatures), R, L, int(color), coefficients(betas),
              float(K), float(delta), float(g_d))
