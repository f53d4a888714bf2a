"""Kawamura, PRL 82, 964 (1999), Eq. (3) (models/kawamura/model.py of the reference).

H = -J sum_<ij> (1 + tau_i tau_j) cos(theta_i - theta_j) (d_i . d_j).
State components per site: theta, tau, dx, dy, dz with d a unit vector.
"""

import numpy as np

from core._native import CHAIN_KAWAMURA, ChainSpec, unmodified
from core.model import Model
from mcgpu import dtype_code
from models._device import require_device

from . import measure as _measure
from .measure import MEASUREMENT_NAMES
from .update import metropolis_color


class KawamuraModel(Model):

    def __init__(self, J=1.0, delta=1.5, dtype=np.float32, start="ordered"):
        dtype = np.dtype(dtype)
        if dtype not in (np.float32, np.float64):
            raise TypeError("KawamuraModel dtype must be float32 or float64")
        if start not in ("ordered", "random"):
            raise ValueError("start must be 'ordered' or 'random'")
        self.dtype = dtype
        self.beta_dtype = dtype
        self.J = dtype.type(J)
        self.delta = dtype.type(delta)  # phase proposal half-width, <= pi
        self.start = start

    def initialize_state(self, L, replicas, seed):
        state = np.zeros((replicas, L, L, 5), dtype=self.dtype)
        if self.start == "random":
            rng = np.random.default_rng(seed)
            theta = rng.random((replicas, L, L)) * (2 * np.pi) - np.pi
            state[..., 0] = theta
            pi = self.dtype.type(np.pi)
            state[..., 0][state[..., 0] >= pi] -= self.dtype.type(2 * np.pi)
            state[..., 1] = np.where(rng.random((replicas, L, L)) < 0.5, 1, -1)
            d = rng.normal(size=(replicas, L, L, 3))
            state[..., 2:] = d / np.linalg.norm(d, axis=-1, keepdims=True)
            return state
        state[..., 1] = 1  # ordered chirality, theta = 0, d = (0, 0, 1)
        state[..., 4] = 1
        return state

    def update_color(self, state, rng_states, betas, color, blocks, threads):
        metropolis_color(state, rng_states, betas, color, self.J, self.delta)

    def measure(self, state, temperature_indices=None, cached_energies=None):
        require_device(state, "state")
        values = _measure.measure(state, self.J, temperature_indices, cached_energies)
        return dict(zip(MEASUREMENT_NAMES, values))

    def measure_with_energy(self, state, energies, temperature_indices=None):
        return self.measure(state, temperature_indices=temperature_indices,
                            cached_energies=energies)

    def measure_energy(self, state, output):
        require_device(state, "state")
        _measure.energy(state, self.J, output)

    def parameters(self):
        return {"J": self.J, "delta": self.delta}

    def chain_spec(self, state, betas):
        """The native chain of this model (core._native), or None once a method is replaced."""
        if not unmodified(self, KawamuraModel):
            return None
        return ChainSpec(CHAIN_KAWAMURA, dtype_code(self.dtype), (float(self.J), float(self.delta)), (),
                         MEASUREMENT_NAMES, None)
