"""Model plugin interface (core/model.py:6-62 of the reference).

A model owns its Hamiltonian parameters and launches its own GPU kernels
through libmcgpu; the Runtime owns the device state, RNG streams and betas.
"""

from abc import ABC, abstractmethod

import numpy as np


class Model(ABC):

    @abstractmethod
    def initialize_state(self, L, replicas, seed) -> np.ndarray:
        """Return the initial host state with shape (replicas, L, L, ...)."""

    @abstractmethod
    def update_color(self, state, rng_states, betas, color, blocks, threads):
        """Launch the Metropolis update of one checkerboard colour on the GPU.

        ``blocks``/``threads`` are the reference's launch shape; the HIP
        kernels choose their own geometry, which never changes results.
        """

    @abstractmethod
    def measure(self, state, temperature_indices=None) -> dict:
        """Reduce the GPU state to named host arrays with a leading replica axis.

        ``temperature_indices`` (int32 device array) gives the destination
        temperature of each physical walker; outputs are scattered on the GPU.
        """

    def measure_energy(self, state, output):
        """Write total energies to a caller-owned float64 device array (R,).

        Energies follow physical walker order and use the same Hamiltonian
        for every temperature.  Nothing is copied to the host and nothing is
        allocated by the caller's side.
        """
        raise NotImplementedError(
            "This model does not provide GPU total-energy measurements"
        )

    def measure_with_energy(self, state, energies, temperature_indices=None):
        """Measure observables reusing fresh walker-ordered GPU energies.

        The default ignores the cache and measures everything.
        """
        return self.measure(state, temperature_indices=temperature_indices)

    def parameters(self) -> dict:
        """Physical parameters recorded in run metadata."""
        return {}
