"""Replica exchange on the GPU for a temperature-independent Hamiltonian.

Drop-in for core/parallel_tempering.py of the reference: the same
constructor, public attributes, methods, diagnostics dictionary, validation
messages and exchange schedule (construction :94-188, sweep :245-249,
attempt_exchanges :251-269, measure :271-273, snapshot :275-278,
diagnostics :280-317).  Configurations and site RNG streams never move: an
exchange swaps walker betas and the two walker/temperature maps on the
device (``mc_pt_exchange``, the counterpart of _exchange_kernel :35-91).

How it is built here:
* ``ExchangeBook``: the swap counters (3, R-1) and round-trip records
  (5, R) share one int64 device allocation, read back by one copy;
* ``WalkerTrail``: the tracked walkers' temperatures in one device buffer
  of rows that doubles when full (device-to-device copy), with the sweep and
  round of every row kept on the host;
* ``advance(n)`` runs n sweeps with their exchanges and trail rows as one
  native call (core._native) when neither this object's nor the runtime's
  sweep and exchange methods are replaced; ``sweep()`` keeps the
  reference's one-sweep-per-call contract.
"""

import operator

import numpy as np

from mcgpu import (as_pointer, create_xoroshiro128p_states, device_array, device_zeros, dtype_code,
                   to_device)
from mcgpu import _lib

ORIGIN, VISITED_OTHER, START_SWEEP, ROUND_TRIPS, ROUND_TRIP_SWEEPS = range(5)


def positive_int(value, name):
    """An int >= 1 (bools refused), or ValueError naming the argument."""
    bad = ValueError(f"{name} must be a positive integer")
    if isinstance(value, (bool, np.bool_)):
        raise bad
    try:
        value = operator.index(value)
    except TypeError:
        raise bad from None
    if value < 1:
        raise bad
    return value


def tracked_ids(tracked_walkers, R):
    """Distinct walker IDs in range(R) as int32, or ValueError."""
    message = "tracked_walkers must contain integer walker IDs"
    try:
        values = list(tracked_walkers)
        if any(isinstance(v, (bool, np.bool_)) for v in values):
            raise ValueError(message)
        ids = np.asarray([operator.index(v) for v in values], dtype=np.int64)
    except (TypeError, OverflowError):
        raise ValueError(message) from None
    if ids.ndim != 1 or ids.size == 0 or ids.min() < 0 or ids.max() >= R or np.unique(ids).size != ids.size:
        raise ValueError("tracked_walkers must be distinct IDs in range(R)")
    return ids.astype(np.int32)


def ladder_betas(runtime):
    """The runtime's betas, which must still be its temperature ladder, strictly decreasing."""
    betas = runtime.betas.copy_to_host()
    if not np.array_equal(betas, (1.0 / runtime.temperatures).astype(betas.dtype)):
        raise ValueError("Runtime betas must initially match its temperature order")
    if not (np.all(np.isfinite(betas) & (betas > 0)) and np.all(betas[1:] < betas[:-1])):
        raise ValueError("PT temperatures must have distinct, finite positive betas in the update precision")
    return betas


class ExchangeBook:
    """Swap counters and round-trip records of R walkers in one int64 device allocation."""

    def __init__(self, R):
        self.R = R
        pairs = R - 1
        start = np.zeros((5, R), dtype=np.int64)
        start[ORIGIN] = -1
        start[START_SWEEP] = -1
        start[ORIGIN, 0], start[ORIGIN, R - 1] = 0, R - 1  # the endpoints start as their own origin
        start[START_SWEEP, [0, R - 1]] = 0
        self.store = to_device(np.concatenate([np.zeros(3 * pairs, dtype=np.int64), start.ravel()]))
        self.counts = self.store[:3 * pairs]   # (3, R-1): attempted, accepted, nonfinite
        self.trips = self.store[3 * pairs:]    # (5, R)

    def read(self):
        host = self.store.copy_to_host()
        pairs = self.R - 1
        return host[:3 * pairs].reshape(3, pairs), host[3 * pairs:].reshape(5, self.R)


class WalkerTrail:
    """Temperature indices of tracked walkers, one device row per recording."""

    def __init__(self, walker_ids, stride):
        self.ids = walker_ids
        self.stride = stride
        self.ids_device = to_device(walker_ids)
        self.rows = device_array((64, walker_ids.size), dtype=np.int32)
        self.used = 0
        self.sweeps = []
        self.rounds = []

    def reserve(self, extra):
        """Room for `extra` more rows (the buffer doubles, keeping what it holds)."""
        need = self.used + extra
        if need <= self.rows.shape[0]:
            return
        size = self.rows.shape[0]
        while size < need:
            size *= 2
        grown = device_array((size, self.ids.size), dtype=np.int32)
        if self.used:
            grown[:self.used].copy_to_device(self.rows[:self.used])
        self.rows = grown

    def record(self, w2t, sweep, rounds):
        self.reserve(1)
        _lib.call("mc_pt_record_walkers", None, as_pointer(w2t), as_pointer(self.ids_device), self.ids.size,
                  as_pointer(self.rows[self.used]))
        self.used += 1
        self.sweeps.append(sweep)
        self.rounds.append(rounds)

    def export(self, w2t_host, sweeps, rounds):
        rows = self.rows[:self.used].copy_to_host() if self.used else np.zeros((0, self.ids.size), np.int32)
        at_sweeps = np.asarray(self.sweeps, dtype=np.int64)
        at_rounds = np.asarray(self.rounds, dtype=np.int64)
        if not self.used or at_rounds[-1] != rounds or at_sweeps[-1] != sweeps:
            # the current positions close the trail (nothing is recorded)
            rows = np.concatenate([rows, w2t_host[self.ids][None, :]])
            at_sweeps = np.append(at_sweeps, sweeps)
            at_rounds = np.append(at_rounds, rounds)
        return {
            "walker_ids": self.ids.copy(),
            "sweeps": at_sweeps,
            "exchange_rounds": at_rounds,
            "temperature_indices": rows,
            "stride": self.stride,
            "stride_units": "exchange_rounds",
        }


class ParallelTempering:
    """Wrap a Runtime; pass this object to Simulation.

    Temperatures must be strictly increasing.  A sweep runs the four colours
    before any exchange.  Matchings alternate between pairs (0,1),(2,3),...
    and (1,2),(3,4),...; with two walkers the single pair is tried every
    exchange interval.  Each matching satisfies detailed balance.

    Diagnostics accumulate from construction (thermalisation included).
    ``tracked_walkers`` records those walkers' temperature indices on the
    device every ``history_stride`` exchange rounds.
    """

    def __init__(self, runtime, exchange_interval=10, seed=None, tracked_walkers=None, history_stride=10):
        interval = positive_int(exchange_interval, "exchange_interval")
        if runtime.R < 2:
            raise ValueError("Parallel tempering requires at least two temperatures")
        stride = positive_int(history_stride, "history_stride")
        ids = None if tracked_walkers is None else tracked_ids(tracked_walkers, runtime.R)
        betas = ladder_betas(runtime)

        self.runtime = runtime
        self.model = runtime.model
        self.L = runtime.L
        self.R = runtime.R
        self.temperatures = runtime.temperatures
        self.seed = runtime.seed
        self.exchange_interval = interval
        self.exchange_seed = operator.index(runtime.seed + 1 if seed is None else seed)
        self.sweeps = 0
        self.exchange_rounds = 0
        self._parity = 0
        self._ladder = betas
        self._beta_code = dtype_code(betas.dtype)

        self.energies = device_array(self.R, dtype=np.float64)
        runtime.measure_energy(self.energies)  # fails here when the model has no GPU energies
        maps = to_device(np.tile(np.arange(self.R, dtype=np.int32), (2, 1)))
        self._maps = maps
        self.walker_to_temperature = maps[0]
        self.temperature_to_walker = maps[1]
        self._pt_streams = create_xoroshiro128p_states(self.R - 1, seed=self.exchange_seed)
        self._book = ExchangeBook(self.R)
        self._trail = None
        if ids is not None:
            self._trail = WalkerTrail(ids, stride)
            self._trail.record(self.walker_to_temperature, 0, 0)

    # -------------------------------------------------------------- schedule
    def sweep(self):
        self.runtime.sweep()
        self.sweeps += 1
        if self.sweeps % self.exchange_interval == 0:
            self.attempt_exchanges()

    def attempt_exchanges(self):
        """Fresh total energies, then the next matching, all on the device."""
        self.runtime.measure_energy(self.energies)
        _lib.call("mc_pt_exchange", None, self._beta_code, as_pointer(self.energies),
                  as_pointer(self.runtime.betas), as_pointer(self.walker_to_temperature),
                  as_pointer(self.temperature_to_walker), as_pointer(self._pt_streams),
                  as_pointer(self._book.counts), as_pointer(self._book.trips), self.R, self._parity,
                  self.sweeps)
        self._after_rounds(1)
        if self._trail is not None and self.exchange_rounds % self._trail.stride == 0:
            self._trail.record(self.walker_to_temperature, self.sweeps, self.exchange_rounds)

    def _after_rounds(self, n):
        self.exchange_rounds += n
        if self.R > 2 and n % 2:
            self._parity = 1 - self._parity

    def _native_chain(self):
        """The runtime's native chain when every step of sweep() is the built-in one."""
        mine = vars(self)
        if ("sweep" in mine or "attempt_exchanges" in mine or type(self).sweep is not ParallelTempering.sweep
                or type(self).attempt_exchanges is not ParallelTempering.attempt_exchanges):
            return None
        sweeps_native = getattr(self.runtime, "_sweeps_native", None)
        if sweeps_native is None or not sweeps_native() or "measure_energy" in vars(self.runtime):
            return None
        return self.runtime.native_chain()

    def advance(self, nsweeps):
        """nsweeps calls of sweep(): one native call when the model has a chain."""
        chain = self._native_chain() if nsweeps > 0 else None
        if chain is None:
            for _ in range(nsweeps):
                self.sweep()
            return
        first, interval = self.sweeps, self.exchange_interval
        exchange_sweeps = [s for s in range((first // interval + 1) * interval, first + nsweeps + 1, interval)]
        trail = self._trail
        history = None
        records = []
        if trail is not None:
            records = [(s, self.exchange_rounds + k + 1) for k, s in enumerate(exchange_sweeps)
                       if (self.exchange_rounds + k + 1) % trail.stride == 0]
            trail.reserve(len(records))
            history = (trail.stride, trail.ids_device, trail.rows, trail.rows.shape[0], trail.used)
        chain.attach_exchange(self.energies, self.walker_to_temperature, self.temperature_to_walker,
                              self._pt_streams, self._book.counts, self._book.trips)
        parity, rounds, row = chain.run(nsweeps, first, interval, self._parity, self.exchange_rounds, history)
        stamp = self.runtime._energy
        stamp.advance(nsweeps)
        if exchange_sweeps and exchange_sweeps[-1] == first + nsweeps:
            stamp.record(self.energies)  # the last sweep ended with a fresh energy reduction
        self.sweeps += nsweeps
        self.exchange_rounds = rounds
        self._parity = parity
        if trail is not None:
            for s, r in records:
                trail.sweeps.append(s)
                trail.rounds.append(r)
            trail.used = row

    # --------------------------------------------------------------- outputs
    def measure(self):
        """Host observables in the fixed temperature order."""
        return self.runtime.measure(temperature_indices=self.walker_to_temperature)

    def sample_channels(self):
        """Channel names that measure_into() writes, or None (see Runtime.sample_channels)."""
        channels = getattr(self.runtime, "sample_channels", None)
        return None if channels is None or "measure" in vars(self) else channels()

    def measure_into(self, out):
        """measure() into a (C, R) float64 device row; channel names, or None without a native chain."""
        into = getattr(self.runtime, "measure_into", None)
        return None if into is None else into(out, temperature_indices=self.walker_to_temperature)

    def snapshot(self):
        """Host snapshot in temperature order; the Runtime stays walker ordered."""
        return self.runtime.snapshot()[self.temperature_to_walker.copy_to_host()]

    def diagnostics(self):
        """Copy the small diagnostic arrays to the host."""
        counts, trips = self._book.read()
        if np.any(counts[2]):
            raise FloatingPointError("Nonfinite energies encountered during replica exchange")
        w2t, t2w = self._maps.copy_to_host()
        attempts, accepted = counts[0].copy(), counts[1].copy()
        completed = int(trips[ROUND_TRIPS].sum())
        out = {
            "exchange_interval": self.exchange_interval,
            "exchange_seed": self.exchange_seed,
            "sweeps": self.sweeps,
            "exchange_rounds": self.exchange_rounds,
            "next_parity": self._parity,
            "diagnostics_scope": "since_wrapper_creation_including_thermalization",
            "temperature_betas": self._ladder.copy(),
            "attempts": attempts,
            "accepted": accepted,
            "acceptance_rates": np.divide(accepted, attempts, out=np.zeros(self.R - 1), where=attempts > 0),
            "walker_to_temperature": w2t,
            "temperature_to_walker": t2w,
            "walker_round_trips": trips[ROUND_TRIPS].copy(),
            "walker_round_trip_total_sweeps": trips[ROUND_TRIP_SWEEPS].copy(),
            "mean_round_trip_sweeps": float(trips[ROUND_TRIP_SWEEPS].sum()) / completed if completed else None,
        }
        if self._trail is not None:
            out["walker_history"] = self._trail.export(w2t, self.sweeps, self.exchange_rounds)
        return out
