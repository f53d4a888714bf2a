"""XY colour update (models/xy/update.py:6-54 of the reference) on gfx950."""

from mcgpu import _lib, as_pointer
from models._device import STREAM, check_update_args, real_dtype


def metropolis_color(state, rng_states, betas, color, J, delta):
    """Uniform-window angle proposal theta + delta (2u - 1), single wrap to [-pi, pi)."""
    R, L, code = check_update_args(state, rng_states, betas, color)
    _lib.call("mc_xy_update_color", STREAM, code, as_pointer(state), as_pointer(rng_states),
              as_pointer(betas), R, L, int(color), float(J), float(delta))


def gaussian_color(state, rng_states, betas, color, J, tables):
    """Gaussian proposal theta + sigma_t z (per-temperature sigma_t), reference acceptance."""
    R, L, code = check_update_args(state, rng_states, betas, color)
    if tables.dtype != real_dtype(state.dtype):
        raise TypeError("proposal tables and state have different precisions")
    _lib.call("mc_xy_gauss_update_color", STREAM, code, as_pointer(state), as_pointer(rng_states),
              as_pointer(betas), R, L, int(color), float(J), as_pointer(tables.table_betas),
              as_pointer(tables.sigmas), tables.T, as_pointer(tables.accepted))
