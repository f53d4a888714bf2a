"""2D XY model with angle spins (models/xy/model.py of the reference).

``dtype=np.float32`` is the reference precision (float32 angles, betas and
sweep arithmetic).  ``dtype=np.float64`` runs the same algorithm in float64:
float64 angles and betas and float64 uniforms (53-bit draws).
"""

import math

import numpy as np

from core.model import Model
from models._device import require_device

from . import measure as _measure
from .measure import MEASUREMENT_NAMES
from .update import metropolis_color


def random_angles(shape, seed, dtype):
    """Uniform angles in [-pi, pi) from numpy's default_rng(seed), stored in dtype."""
    dtype = np.dtype(dtype)
    theta = (np.random.default_rng(seed).random(shape) * (2 * np.pi) - np.pi).astype(dtype)
    pi = dtype.type(np.pi)
    theta[theta >= pi] -= dtype.type(2 * np.pi)
    return theta


class XYModel(Model):

    def __init__(self, J=1.0, delta=1.5, dtype=np.float32, start="ordered"):
        if J <= 0:
            raise ValueError("J must be positive")
        if not (0 < delta <= math.pi):
            raise ValueError("delta must be in (0, pi]")
        dtype = np.dtype(dtype)
        if dtype not in (np.float32, np.float64):
            raise TypeError("XYModel dtype must be float32 or float64")
        if start not in ("ordered", "random"):
            raise ValueError("start must be 'ordered' or 'random'")
        self.dtype = dtype
        self.beta_dtype = dtype
        self.J = dtype.type(J)
        self.delta = dtype.type(delta)
        self.start = start

    def initialize_state(self, L, replicas, seed):
        if self.start == "random":
            return random_angles((replicas, L, L), seed, self.dtype)
        # Ordered start, theta = 0 everywhere.
        return np.zeros((replicas, L, L), dtype=self.dtype)

    def update_color(self, state, rng_states, betas, color, blocks, threads):
        metropolis_color(state, rng_states, betas, color, self.J, self.delta)

    def measure(self, state, temperature_indices=None, cached_energies=None):
        """Reduce the GPU state to the five host measurement arrays."""
        require_device(state, "state")
        results = _measure.measure(state, self.J, temperature_indices, cached_energies)
        return {name: results[c] for c, name in enumerate(MEASUREMENT_NAMES)}

    def measure_with_energy(self, state, energies, temperature_indices=None):
        return self.measure(state, temperature_indices=temperature_indices,
                            cached_energies=energies)

    def parameters(self):
        return {"J": self.J, "delta": self.delta}

    def measure_energy(self, state, output):
        """Leave total energies on the GPU for replica exchange."""
        require_device(state, "state")
        _measure.energy(state, self.J, output)
