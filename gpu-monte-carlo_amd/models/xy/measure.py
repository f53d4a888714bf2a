"""XY measurements: energy, mx, my, twist_cos_x, twist_sin_x
(models/xy/measure.py of the reference), float64 sums on the GPU."""

from mcgpu import _lib, as_pointer, dtype_code
from models._device import STREAM, energies_pointer, indices_pointer, output_buffer

MEASURE_THREADS = 256
MEASUREMENT_NAMES = ("energy", "mx", "my", "twist_cos_x", "twist_sin_x")


def measure(state, J, temperature_indices=None, cached_energies=None):
    R, L = state.shape[0], state.shape[1]
    out = output_buffer(len(MEASUREMENT_NAMES), R)
    _lib.call("mc_xy_measure", STREAM, dtype_code(state.dtype), as_pointer(state), R, L, float(J),
              indices_pointer(temperature_indices, R), energies_pointer(cached_energies, R),
              as_pointer(out))
    return out.copy_to_host()


def energy(state, J, output):
    R, L = state.shape[0], state.shape[1]
    _lib.call("mc_xy_energy", STREAM, dtype_code(state.dtype), as_pointer(state), R, L, float(J),
              energies_pointer(output, R))
