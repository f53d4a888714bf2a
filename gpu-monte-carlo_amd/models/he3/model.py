"""Full complex 3x2 thin-film Ginzburg-Landau model (models/he3/model.py of the reference)."""

import numpy as np

from core._native import CHAIN_HE3, ChainSpec, unmodified
from core.model import Model
from mcgpu import dtype_code
from models._device import require_device

from . import measure as _measure
from .measure import MEASUREMENT_NAMES
from .update import metropolis_color

# Strong-coupling corrections: Regan, Wiman & Sauls, PRB 101, 024517 (2020),
# Appendix Table I (arXiv:1908.04190).  Rows are powers 0..5 of p/34 bar,
# columns delta_beta_i / |beta_1^wc|; calibrated for 0 <= p <= 34 bar.
_STRONG_COUPLING = np.array([
    [-0.009849, -0.04193, -0.01322, -0.004747, -0.08987],
    [-0.05043, -0.1177, -0.05428, -0.3788, -0.6925],
    [0.02205, -0.04322, 0.09559, -0.1774, 0.8761],
    [-0.02557, 0.08793, -0.06419, 0.1735, -0.5929],
    [0.05023, -0.08598, -0.009310, 0.1878, 0.02904],
    [-0.02769, 0.03639, 0.01862, -0.1522, 0.08870],
])


class He3Model(Model):

    def __init__(self, K=1.0, betas=(-0.5, 1, 1, 1, -1), delta=0.15, pressure=None,
                 reduced_temperature=1.0, g_d=0, dtype=np.complex64):
        """Pressure corrections to the GL betas at a fixed reduced temperature.

        pressure=None keeps the given betas; pressure=0 adds the zero-bar
        strong-coupling correction.  The lattice coefficients are
        beta_i / (2 |beta_1^wc|), hence the factor 0.5.  ``dtype`` is
        complex64 (reference precision) or complex128 (float64 mode).
        """
        dtype = np.dtype(dtype)
        if dtype not in (np.complex64, np.complex128):
            raise TypeError("He3Model dtype must be complex64 or complex128")
        self.dtype = dtype
        real = np.float32 if dtype == np.complex64 else np.float64
        self.beta_dtype = np.dtype(real)
        self.K = real(K)
        self.pressure = None if pressure is None else float(pressure)
        self.reduced_temperature = float(reduced_temperature)
        coefficients = np.asarray(betas, dtype=np.float64)
        if self.pressure is not None:
            correction = np.polynomial.polynomial.polyval(self.pressure / 34, _STRONG_COUPLING)
            coefficients = coefficients + 0.5 * self.reduced_temperature * correction
        self.betas = tuple(real(b) for b in coefficients)
        self.delta = real(delta)
        if not np.isfinite(g_d) or g_d < 0:
            raise ValueError("g_d must be finite and nonnegative")
        self.g_d = real(g_d)

    def initialize_state(self, L, replicas, seed):
        state = np.zeros((replicas, L, L, 3, 2), dtype=self.dtype)
        # Homogeneous positive-chirality minimum: d = (0, 0, 1), theta = 0.
        beta_chiral = self.betas[1] + self.betas[3] + self.betas[4]
        amplitude = self.beta_dtype.type(1 / np.sqrt(2 * beta_chiral))
        state[..., 2, 0] = amplitude
        state[..., 2, 1] = 1j * amplitude
        return state

    def update_color(self, state, rng_states, betas, color, blocks, threads):
        metropolis_color(state, rng_states, betas, color, self.betas, self.K, self.delta, self.g_d)

    def measure(self, state, temperature_indices=None, cached_energies=None):
        require_device(state, "state")
        values = _measure.measure(state, self.betas, self.K, self.g_d, temperature_indices,
                                  cached_energies)
        return dict(zip(MEASUREMENT_NAMES, values))

    def measure_with_energy(self, state, energies, temperature_indices=None):
        return self.measure(state, temperature_indices=temperature_indices,
                            cached_energies=energies)

    def measure_energy(self, state, output):
        require_device(state, "state")
        _measure.energy(state, self.betas, self.K, self.g_d, output)

    def chain_spec(self, state, betas):
        """The native chain of this model (core._native), or None once a method is replaced."""
        if not unmodified(self, He3Model):
            return None
        params = (*map(float, self.betas), float(self.K), float(self.delta), float(self.g_d))
        return ChainSpec(CHAIN_HE3, dtype_code(self.dtype), params, (), MEASUREMENT_NAMES, None)

    def parameters(self):
        return {"K": self.K, "betas": self.betas, "delta": self.delta,
                "pressure": self.pressure, "reduced_temperature": self.reduced_temperature,
                "g_d": float(self.g_d),
                "dipole_length_lattice": (float(np.sqrt(2.0 * self.K / self.g_d))
                                          if self.g_d > 0 else None)}
