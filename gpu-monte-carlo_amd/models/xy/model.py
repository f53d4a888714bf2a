"""2D XY model with angle spins (models/xy/model.py of the reference).

``dtype=np.float32`` is the reference precision (float32 angles, betas and
sweep arithmetic).  ``dtype=np.float64`` runs the same algorithm in float64:
float64 angles and betas and float64 uniforms (53-bit draws).

``proposal="uniform"`` (default) is the reference's window theta + delta
(2u - 1).  ``proposal="gaussian"`` draws theta + sigma_t z with a width per
ladder temperature (models/xy/proposals.py); a Runtime attaches the tables.
"""

import math

import numpy as np

from core._native import CHAIN_XY, CHAIN_XY_GAUSS, ChainSpec, unmodified
from core.model import Model
from mcgpu import dtype_code
from models._device import require_device

from . import measure as _measure
from .measure import MEASUREMENT_NAMES
from .proposals import GaussianTables, ProposalRegistry
from .update import gaussian_color, metropolis_color


def random_angles(shape, seed, dtype):
    """Uniform angles in [-pi, pi) from numpy's default_rng(seed), stored in dtype."""
    dtype = np.dtype(dtype)
    theta = (np.random.default_rng(seed).random(shape) * (2 * np.pi) - np.pi).astype(dtype)
    pi = dtype.type(np.pi)
    theta[theta >= pi] -= dtype.type(2 * np.pi)
    return theta


class XYModel(Model):

    def __init__(self, J=1.0, delta=1.5, dtype=np.float32, start="ordered", proposal="uniform", sigma=1.0):
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
        if proposal not in ("uniform", "gaussian"):
            raise ValueError("proposal must be 'uniform' or 'gaussian'")
        sig = np.asarray(sigma, dtype=np.float64)
        if not np.all((sig > 0) & np.isfinite(sig)):
            raise ValueError("sigma must be positive and finite")
        self.proposal = proposal
        self.sigma = sig
        self._proposals = ProposalRegistry()

    def attach(self, runtime):
        """Gaussian proposals: per-temperature width tables for this runtime's ladder."""
        if self.proposal == "gaussian":
            self._proposals.attach(runtime.betas, GaussianTables(runtime.temperatures, runtime.L, self.sigma,
                                                                 self.dtype))

    def proposal_tables(self, runtime):
        """Width tables and acceptance counters of a runtime (Gaussian proposals)."""
        return self._proposals.get(runtime.betas)

    def initialize_state(self, L, replicas, seed):
        if self.start == "random":
            return random_angles((replicas, L, L), seed, self.dtype)
        # Ordered start, theta = 0 everywhere.
        return np.zeros((replicas, L, L), dtype=self.dtype)

    def update_color(self, state, rng_states, betas, color, blocks, threads):
        if self.proposal == "gaussian":
            tables = self._proposals.get(betas)
            gaussian_color(state, rng_states, betas, color, self.J, tables)
            tables.count_colour()
        else:
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
        if self.proposal == "gaussian":
            return {"J": self.J, "proposal": "gaussian", "sigma": self.sigma.tolist()}
        return {"J": self.J, "delta": self.delta}

    def measure_energy(self, state, output):
        """Leave total energies on the GPU for replica exchange."""
        require_device(state, "state")
        _measure.energy(state, self.J, output)

    def chain_spec(self, state, betas):
        """The native chain of this model (core._native), or None once a method is replaced."""
        if not unmodified(self, XYModel):
            return None
        code = dtype_code(self.dtype)
        if self.proposal == "gaussian":
            tables = self._proposals.get(betas)
            return ChainSpec(CHAIN_XY_GAUSS, code, (float(self.J), float(tables.T)),
                             (tables.table_betas, tables.sigmas, tables.accepted), MEASUREMENT_NAMES,
                             tables.count_sweeps)
        return ChainSpec(CHAIN_XY, code, (float(self.J), float(self.delta)), (), MEASUREMENT_NAMES, None)
