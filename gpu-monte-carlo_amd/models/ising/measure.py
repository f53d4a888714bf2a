"""Ising energy and magnetisation sums (models/ising/measure.py of the reference)."""

from mcgpu import _lib, as_pointer
from models._device import STREAM, energies_pointer, indices_pointer, output_buffer

MEASURE_THREADS = 256
MEASUREMENT_NAMES = ("energy", "magnetization")


def measure(spins, J, temperature_indices=None, cached_energies=None):
    """Return the (2, R) float64 host array of channel sums."""
    R, L = spins.shape[0], spins.shape[1]
    out = output_buffer(len(MEASUREMENT_NAMES), R)
    _lib.call("mc_ising_measure", STREAM, as_pointer(spins), R, L, float(J),
              indices_pointer(temperature_indices, R), energies_pointer(cached_energies, R),
              as_pointer(out))
    return out.copy_to_host()


def energy(spins, J, output):
    R, L = spins.shape[0], spins.shape[1]
    _lib.call("mc_ising_energy", STREAM, as_pointer(spins), R, L, float(J),
              energies_pointer(output, R))
