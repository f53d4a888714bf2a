"""Native chains: the sweep / exchange loop of the drop-in path in one C call.

A built-in model describes itself to the library through ``chain_spec``
(the kind, precision, Hamiltonian parameters and extra device tables of an
``mc_chain``); ``NativeChain`` binds that description to a runtime's state,
RNG streams and betas.  The drop-in classes use a chain whenever the model's
behaviour is the built-in one -- no method replaced on the instance or in a
subclass -- and otherwise call the model's methods one launch at a time, as
the reference does (core/runtime.py:87-99).
"""

import ctypes
import sys
from collections import namedtuple

from mcgpu import _lib, as_pointer

CHAIN_ISING, CHAIN_XY, CHAIN_XY_GAUSS, CHAIN_KAWAMURA, CHAIN_HE3, CHAIN_HE3_GAUSS = range(6)

# kind, precision code, float parameters, extra device buffers (kept alive by
# the spec), measurement channel names, and a callback run after n sweeps
# (host-side counters that the per-colour path keeps)
ChainSpec = namedtuple("ChainSpec", "kind code params buffers names after_sweeps")

MODEL_METHODS = ("update_color", "measure", "measure_with_energy", "measure_energy")


def unmodified(obj, cls, names=MODEL_METHODS):
    """True when obj behaves as cls for `names`: none is set on the instance
    or replaced by a subclass."""
    inst = vars(obj) if hasattr(obj, "__dict__") else {}
    return all(n not in inst and getattr(type(obj), n, None) is getattr(cls, n) for n in names)


class NativeChain:
    """An ``mc_chain`` bound to one runtime's device buffers."""

    def __init__(self, spec, state, rng_states, betas, L, R):
        self.spec = spec
        self.L, self.R = L, R
        params = (ctypes.c_double * max(len(spec.params), 1))(*spec.params)
        bufs = (ctypes.c_void_p * max(len(spec.buffers), 1))(*[as_pointer(b) for b in spec.buffers])
        handle = ctypes.c_void_p()
        _lib.call("mc_chain_create", spec.kind, spec.code, L, R, as_pointer(state), as_pointer(rng_states),
                  as_pointer(betas), params, len(spec.params), bufs, len(spec.buffers), ctypes.byref(handle))
        self._handle = handle
        self._exchange = None

    def __del__(self, _finalizing=sys.is_finalizing):
        h = getattr(self, "_handle", None)
        if h is not None and h.value and not _finalizing() and not _lib.closing():
            _lib.load().mc_chain_destroy(h)

    @property
    def names(self):
        return self.spec.names

    def sweeps(self, n, stream=None):
        _lib.call("mc_chain_sweeps", self._handle, stream, int(n))
        if self.spec.after_sweeps is not None:
            self.spec.after_sweeps(int(n))

    def energy(self, out, stream=None):
        _lib.call("mc_chain_energy", self._handle, stream, as_pointer(out))

    def measure(self, out, temperature_indices=None, cached_energies=None, stream=None):
        _lib.call("mc_chain_measure", self._handle, stream, as_pointer(temperature_indices),
                  as_pointer(cached_energies), as_pointer(out))

    def attach_exchange(self, energies, w2t, t2w, pt_rng, counts, trips):
        key = tuple(a.ptr for a in (energies, w2t, t2w, pt_rng, counts, trips))
        if key != self._exchange:
            _lib.call("mc_chain_attach_exchange", self._handle, *[as_pointer(a) for a in
                                                                  (energies, w2t, t2w, pt_rng, counts, trips)])
            self._exchange = key

    def run(self, nsweeps, sweep0, interval, parity, rounds0, history=None, stream=None):
        """nsweeps sweeps with exchanges; history = (stride, tracked, buffer, capacity, row0) or None.
        Returns (parity, rounds, next history row)."""
        par, rounds, row = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong()
        if history is None:
            stride, tracked, buf, cap, row0, ntracked = 1, None, None, 0, 0, 0
        else:
            stride, tracked, buf, cap, row0 = history
            ntracked = tracked.size
        _lib.call("mc_chain_run", self._handle, stream, int(nsweeps), int(sweep0), int(interval), int(parity),
                  int(rounds0), int(stride), as_pointer(tracked), int(ntracked), as_pointer(buf), int(cap),
                  int(row0), ctypes.byref(par), ctypes.byref(rounds), ctypes.byref(row))
        if self.spec.after_sweeps is not None:
            self.spec.after_sweeps(int(nsweeps))
        return int(par.value), int(rounds.value), int(row.value)
