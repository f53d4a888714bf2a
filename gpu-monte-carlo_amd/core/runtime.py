"""Runtime: the walkers of one lattice size on the GPU.

Drop-in for core/runtime.py of the reference: the same constructor, public
attributes, methods, validation messages and energy-cache rules
(construction :10-73, sweep :87-99, measure :101-119, measure_energy
:121-132, snapshot :134-135).

How it is built here:
* device buffers (state, RNG streams, betas) are plain ``mcgpu.DeviceArray``
  objects that callers may read, overwrite or slice;
* a sweep is one call into the library: the model's ``chain_spec`` names a
  native chain (core._native) that queues the four colour launches from C.
  A model without one, or whose methods were replaced, is swept through
  ``update_color`` one colour at a time, as the plugin contract says;
* the energy cache is an ``_EnergyStamp``: a configuration epoch that every
  sweep and every explicit invalidation advances, and the caller's buffer
  together with the epoch its energies belong to.
"""

import numpy as np

from mcgpu import create_xoroshiro128p_states, to_device


def checked_ladder(temperatures):
    """The temperatures as a fresh float64 vector, or ValueError."""
    ladder = np.array(temperatures, dtype=np.float64)
    if ladder.ndim != 1 or ladder.size == 0:
        raise ValueError("Temperatures must be a nonempty 1D array")
    if not np.all(np.isfinite(ladder) & (ladder > 0)):
        raise ValueError("Temperatures must be positive and finite")
    return ladder


def checked_geometry(L, threads):
    if L < 4 or L % 2 != 0:
        raise ValueError("L must be even and >= 4")
    if threads < 1 or threads**2 > 1024:
        raise ValueError("Invalid thread configuration")


def launch_shape(L, R, threads):
    """The reference's per-colour launch geometry (one thread per site of a colour)."""
    blocks = -(-(L // 2) // threads)
    return (threads, threads, 1), (blocks, blocks, R)


class _EnergyStamp:
    """Walker energies kept in a caller's device buffer, valid for one configuration epoch."""

    __slots__ = ("epoch", "buffer", "stamp")

    def __init__(self):
        self.epoch = 0
        self.buffer = None
        self.stamp = -1

    def advance(self, n=1):
        """The configuration changes (n sweeps or an external edit): the energies expire."""
        self.epoch += n
        self.forget()

    def forget(self):
        self.buffer = None
        self.stamp = -1

    def record(self, buffer):
        self.buffer = buffer
        self.stamp = self.epoch

    def current(self):
        return self.buffer if self.buffer is not None and self.stamp == self.epoch else None


class Runtime:

    def __init__(self, model, L, temperatures, threads=16, seed=42):
        checked_geometry(L, threads)
        ladder = checked_ladder(temperatures)
        R = ladder.size

        host = np.asarray(model.initialize_state(L, R, seed))
        if host.shape[:3] != (R, L, L):
            raise ValueError("Invalid model state shape")

        self.model = model
        self.L = L
        self.R = R
        self.temperatures = ladder
        self.seed = seed
        self.state = to_device(np.ascontiguousarray(host))
        self.rng_states = create_xoroshiro128p_states(R * L * L, seed=seed)
        self.betas = to_device((1.0 / ladder).astype(np.dtype(getattr(model, "beta_dtype", np.float32))))
        bind = getattr(model, "attach", None)  # per-ladder device tables (Gaussian XY widths)
        if bind is not None:
            bind(self)
        self.threads_per_block, self.blocks_per_grid = launch_shape(L, R, threads)

        self._energy = _EnergyStamp()
        self._chain = None
        self._chain_key = None

    # ------------------------------------------------------------ native chain
    def _chain_spec(self):
        make = getattr(self.model, "chain_spec", None)
        return None if make is None else make(self.state, self.betas)

    def native_chain(self):
        """The native chain bound to the current buffers and model parameters, or None
        when the model has none (or a replaced method)."""
        spec = self._chain_spec()
        if spec is None:
            return None
        key = (spec.kind, spec.code, spec.params, tuple(b.ptr for b in spec.buffers),
               self.state.ptr, self.rng_states.ptr, self.betas.ptr)
        if key != self._chain_key:
            from core._native import NativeChain
            self._chain = NativeChain(spec, self.state, self.rng_states, self.betas, self.L, self.R)
            self._chain_key = key
        else:
            self._chain.spec = spec
        return self._chain

    def _sweeps_native(self):
        """True when advance() may run sweeps without calling self.sweep (not replaced)."""
        return "sweep" not in vars(self) and type(self).sweep is Runtime.sweep

    # ------------------------------------------------------------------ API
    def invalidate_energy_cache(self):
        """Invalidate energies before changing fields or Hamiltonian parameters.

        sweep() does this itself.  Call it after restoring or editing the
        state or changing model parameters.  Temperature exchanges keep
        configurations and need no invalidation.
        """
        self._energy.advance()

    def sweep(self):
        # the stamp expires before any update is queued, so a sweep that
        # fails half way cannot leave energies of an older state in use
        self._energy.advance()
        chain = self.native_chain()
        if chain is not None:
            chain.sweeps(1)
            return
        for color in range(4):
            self.model.update_color(self.state, self.rng_states, self.betas, color,
                                    self.blocks_per_grid, self.threads_per_block)

    def advance(self, nsweeps):
        """nsweeps calls of sweep(): one native call when the model has a chain."""
        chain = self.native_chain() if self._sweeps_native() else None
        if chain is None:
            for _ in range(nsweeps):
                self.sweep()
            return
        self._energy.advance(nsweeps)
        chain.sweeps(nsweeps)

    def measure(self, temperature_indices=None):
        """Host observables, optionally scattered into temperature order."""
        energies = self._energy.current()
        if energies is not None:
            return self.model.measure_with_energy(self.state, energies, temperature_indices=temperature_indices)
        return self.model.measure(self.state, temperature_indices=temperature_indices)

    def sample_channels(self):
        """Channel names that measure_into() writes, or None when it cannot (no native
        chain, or measure() replaced on this object)."""
        if "measure" in vars(self):
            return None
        chain = self.native_chain()
        return None if chain is None else chain.names

    def measure_into(self, out, temperature_indices=None):
        """measure() into row `out` ((C, R) float64 device array) without a host copy;
        the channel names, or None when the model has no native chain."""
        chain = self.native_chain()
        if chain is None:
            return None
        chain.measure(out, temperature_indices, self._energy.current())
        return chain.names

    def measure_energy(self, output):
        """Total energies in physical walker order into a caller-owned buffer."""
        # the previous buffer may be the one being overwritten: it stops
        # counting as fresh before the producer runs
        self._energy.forget()
        result = self.model.measure_energy(self.state, output)
        self._energy.record(output)
        return result

    def snapshot(self):
        return self.state.copy_to_host()
