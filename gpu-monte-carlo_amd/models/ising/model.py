"""2D Ising ferromagnet with int8 spins (models/ising/model.py of the reference)."""

import numpy as np

from core._native import CHAIN_ISING, ChainSpec, unmodified
from core.model import Model
from mcgpu import MC_F32
from models._device import require_device

from . import measure as _measure
from .measure import MEASUREMENT_NAMES
from .update import metropolis_color


class IsingModel(Model):
    """H = -J sum_<ij> s_i s_j on the periodic square lattice.

    ``start="cold"`` (the reference's default) starts every spin at +1;
    ``start="random"`` draws +-1 spins from numpy's default_rng(seed).
    """

    beta_dtype = np.float32

    def __init__(self, J=1.0, start="cold"):
        if start not in ("cold", "random"):
            raise ValueError("start must be 'cold' or 'random'")
        self.J = np.float32(J)
        self.start = start

    def initialize_state(self, L, replicas, seed):
        if self.start == "random":
            draws = np.random.default_rng(seed).random((replicas, L, L))
            return np.where(draws < 0.5, 1, -1).astype(np.int8)
        return np.ones((replicas, L, L), dtype=np.int8)

    def update_color(self, state, rng_states, betas, color, blocks, threads):
        metropolis_color(state, rng_states, betas, color, self.J)

    def measure(self, state, temperature_indices=None, cached_energies=None):
        """Reduce the GPU state to the two host measurement arrays."""
        require_device(state, "state")
        results = _measure.measure(state, self.J, temperature_indices, cached_energies)
        return {name: results[c] for c, name in enumerate(MEASUREMENT_NAMES)}

    def measure_with_energy(self, state, energies, temperature_indices=None):
        return self.measure(state, temperature_indices=temperature_indices,
                            cached_energies=energies)

    def parameters(self):
        return {"J": self.J}

    def chain_spec(self, state, betas):
        """The native chain of this model (core._native), or None once a method is replaced."""
        if not unmodified(self, IsingModel):
            return None
        return ChainSpec(CHAIN_ISING, MC_F32, (float(self.J),), (), MEASUREMENT_NAMES, None)

    def measure_energy(self, state, output):
        """Leave total energies on the GPU for replica exchange."""
        require_device(state, "state")
        _measure.energy(state, self.J, output)
