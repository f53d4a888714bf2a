"""Kawamura colour update (models/kawamura/update.py:21-65 of the reference)."""

from mcgpu import _lib, as_pointer
from models._device import STREAM, check_update_args


def metropolis_color(state, rng_states, betas, color, J, delta):
    """theta move, tau flip, then a uniform-sphere d move, each its own Metropolis test."""
    R, L, code = check_update_args(state, rng_states, betas, color)
    if state.ndim != 4 or state.shape[3] != 5:
        raise ValueError("Kawamura state must have shape (R, L, L, 5)")
    _lib.call("mc_kawamura_update_color", STREAM, code, as_pointer(state), as_pointer(rng_states),
              as_pointer(betas), R, L, int(color), float(J), float(delta))
