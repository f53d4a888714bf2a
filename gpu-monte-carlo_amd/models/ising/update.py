"""Ising colour update (models/ising/update.py:7-29 of the reference) on gfx950."""

from mcgpu import _lib, as_pointer
from models._device import STREAM, check_update_args


def metropolis_color(spins, rng_states, betas, color, J):
    """Flip int8 spins of one colour with float32 Metropolis acceptance."""
    R, L, _ = check_update_args(spins, rng_states, betas, color)
    if spins.dtype.itemsize != 1:
        raise TypeError("Ising spins must be int8")
    _lib.call("mc_ising_update_color", STREAM, as_pointer(spins), as_pointer(rng_states),
              as_pointer(betas), R, L, int(color), float(J))
