"""He3 energy, density, chirality, charge-two phase, exact twist derivatives and
spin alignment (models/he3/measure.py of the reference).

Local algebra uses the state precision; lattice sums and cached energies are
float64.  Qzz = (3 * total spin-z density / total density - 1) / 2, and 0 for
zero total density.
"""

from mcgpu import _lib, as_pointer, dtype_code
from models._device import STREAM, energies_pointer, indices_pointer, output_buffer

from .update import coefficients

MEASURE_THREADS = 256
MEASUREMENT_NAMES = (
    "energy", "density", "chirality", "phase_real", "phase_imag",
    "current_x", "curvature_x", "current_y", "curvature_y",
    "spin_alignment",
)


def measure(state, betas, K, g_d, temperature_indices=None, cached_energies=None):
    R, L = state.shape[0], state.shape[1]
    out = output_buffer(len(MEASUREMENT_NAMES), R)
    _lib.call("mc_he3_measure", STREAM, dtype_code(state.dtype), as_pointer(state), R, L,
              coefficients(betas), float(K), float(g_d), indices_pointer(temperature_indices, R),
              energies_pointer(cached_energies, R), as_pointer(out))
    return out.copy_to_host()


def energy(state, betas, K, g_d, output):
    R, L = state.shape[0], state.shape[1]
    _lib.call("mc_he3_energy", STREAM, dtype_code(state.dtype), as_pointer(state), R, L,
              coefficients(betas), float(K), float(g_d), energies_pointer(output, R))
