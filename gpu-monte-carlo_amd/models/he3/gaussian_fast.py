"""Conditional-Gaussian proposals with a full-Hamiltonian MH correction
(models/he3/gaussian_fast.py of the reference) on gfx950.

Around the conditional mean of the fixed neighbours, every Cartesian field
component gets an AR(rho) Gaussian proposal (orbital eigenvalues 10K+m and
6K+m).  The acceptance uses the full onsite + gradient change and the
reverse/forward proposal density ratio.  Per-site attempt/accept/nonfinite
counters are int64 (3, R, L, L).
"""

import numpy as np

from core._native import CHAIN_HE3_GAUSS, ChainSpec, unmodified
from mcgpu import _lib, as_pointer, device_zeros, dtype_code
from models._device import STREAM

from .measure import MEASUREMENT_NAMES
from .model import He3Model
from .update import check_he3_state, coefficients


class FastGaussianUpdater:
    """step() replaces one sweep of the Runtime it wraps."""

    def __init__(self, runtime, rho=0.0, mass=0.0):
        if len(runtime.state.shape) != 5 or runtime.state.shape[-2:] != (3, 2):
            raise ValueError("FastGaussianUpdater requires full complex 3x2 He3 fields")
        real = runtime.model.beta_dtype.type
        rho, mass = real(rho), real(mass)
        eigenvalues = np.asarray([6, 10], dtype=runtime.model.beta_dtype) * runtime.model.K + mass
        if not np.isfinite(mass) or not np.all(np.isfinite(eigenvalues) & (eigenvalues > 0)):
            raise ValueError("Conditional quadratic Gaussian requires positive finite orbital eigenvalues")
        if not np.isfinite(rho) or not -1 < rho < 1:
            raise ValueError("Require -1 < rho < 1 for nondegenerate proposals")
        self.runtime, self.rho, self.mass = runtime, rho, mass
        self.counts = device_zeros((3, runtime.R, runtime.L, runtime.L), dtype=np.int64)
        self.sweeps = 0

    def chain_spec(self):
        """The native chain of this sampler (core._native), or None once the model's
        measurements are replaced."""
        model = self.runtime.model
        if not unmodified(model, He3Model, ("measure", "measure_with_energy", "measure_energy")):
            return None
        params = (*map(float, model.betas), float(model.K), float(self.rho), float(self.mass), float(model.g_d))
        return ChainSpec(CHAIN_HE3_GAUSS, dtype_code(model.dtype), params, (self.counts,), MEASUREMENT_NAMES,
                         self._count)

    def _count(self, n):
        self.sweeps += n

    def step(self):
        rt = self.runtime
        rt.invalidate_energy_cache()
        check_he3_state(rt.state)
        chain = rt.native_chain()
        if chain is not None:
            chain.sweeps(1)
            return
        model = rt.model
        code = 0 if rt.state.dtype == np.complex64 else 1
        coef = coefficients(model.betas)
        for color in range(4):
            _lib.call("mc_he3_gaussian_update_color", STREAM, code, as_pointer(rt.state),
                      as_pointer(rt.rng_states), as_pointer(rt.betas), rt.R, rt.L, color, coef,
                      float(model.K), float(self.rho), float(self.mass), as_pointer(self.counts),
                      float(model.g_d))
        self.sweeps += 1

    def diagnostics(self):
        counts = self.counts.copy_to_host().sum(axis=(2, 3))
        return {
            "algorithm": "float32_conditional_gaussian_full_hamiltonian_mh"
            if self.runtime.state.dtype == np.complex64
            else "float64_conditional_gaussian_full_hamiltonian_mh",
            "sweeps_including_compilation": self.sweeps,
            "rho": float(self.rho), "proposal_mass": float(self.mass),
            "attempts": counts[0].tolist(), "accepted": counts[1].tolist(),
            "nonfinite_rejections": counts[2].tolist(),
            "acceptance_rates": np.divide(counts[1], counts[0], out=np.zeros(self.runtime.R),
                                          where=counts[0] > 0).tolist(),
        }
