"""XY colour update (models/xy/update.py:6-54 of the reference) on gfx950."""

from mcgpu import _lib, as_pointer
from models._device import STREAM, check_update_args


def metropolis_color(state, rng_states, betas, color, J, delta):
    """Uniform-window angle proposal theta + delta (2u - 1), single wrap to [-pi, pi)."""
    R, L, code = check_update_args(state, rng_states, betas, color)
    _lib.call("mc_xy_update_color", STREAM, code, as_pointer(state), as_pointer(rng_states),
              as_pointer(betas), R, L, int(color), float(J), float(delta))
