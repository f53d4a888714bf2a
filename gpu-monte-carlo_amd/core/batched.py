"""Batched persistent engine: whole runs in one persistent launch per lattice size.

The drop-in path (Runtime + ParallelTempering + Simulation) launches four
colour kernels per sweep from Python, like the reference.  ``BatchedRun``
executes the same schedule -- sweeps (core/runtime.py:87-99), an exchange
every ``exchange_interval`` sweeps (core/parallel_tempering.py:245-269) and a
sample every ``sample_interval`` sweeps (core/simulation.py:36-72) -- inside
one persistent launch (``mc_batch_run``; a run makes a launch plan once with
``mc_batch_plan_create`` and queues each launch with ``mc_batch_plan_launch``):
each workgroup keeps one strip of H rows of one replica on chip for the whole
launch.

Models:
* ``IsingModel``: the reference update; trajectories, exchanges and samples
  are bit-identical to the drop-in path with the same seeds.
* ``XYModel(dtype=float32)``: the float32 fast update (dE from cached cos/sin
  of every site, float32 sincos/exp); energies are exact integer sums of the
  float32 bond terms, each rounded to a multiple of 2^-40, so they do not
  depend on how sites are spread over threads, strips and workgroups.
* ``XYModel(dtype=float32, proposal="gaussian")``: the same with Gaussian
  proposals of per-temperature width; ``adapt_sigmas`` tunes the widths
  toward a target acceptance during thermalisation.
Accepted moves are counted per ladder temperature for every model
(``acceptance_by_temperature``).

``BatchedSimulation`` runs several lattice sizes (a finite-size-scaling set)
concurrently, one stream each, and returns the reference's run dictionaries
(``{L: {"samples", "metadata"}}``, loadable by utils.data_io).
"""

import ctypes
import sys
import operator
import time

import numpy as np

from mcgpu import (PinnedArray, _lib, as_pointer, create_xoroshiro128p_states, device_array, device_zeros,
                   synchronize, to_device)

MODEL_ISING, MODEL_XY_FAST, MODEL_XY_GAUSS = 0, 1, 2
CHANNELS = 5
NAMES = {
    MODEL_ISING: ("energy", "magnetization"),
    MODEL_XY_FAST: ("energy", "mx", "my", "twist_cos_x", "twist_sin_x"),
    MODEL_XY_GAUSS: ("energy", "mx", "my", "twist_cos_x", "twist_sin_x"),
}

_tuned = {}


def model_code(model):
    from models.ising.model import IsingModel
    from models.xy.model import XYModel
    if isinstance(model, IsingModel):
        return MODEL_ISING
    if isinstance(model, XYModel):
        if model.dtype != np.float32:
            raise TypeError("the batched XY path runs float32 sweeps; use XYModel(dtype=np.float32)")
        return MODEL_XY_GAUSS if getattr(model, "proposal", "uniform") == "gaussian" else MODEL_XY_FAST
    raise TypeError(f"no batched kernel for {type(model).__name__}")


WORKGROUP_SIZES = (256, 512)


def max_workgroups(code, L, R, H, exchange, threads=256):
    n = ctypes.c_int(0)
    _lib.call("mc_batch_max_workgroups", code, L, R, H, int(threads), int(bool(exchange)), ctypes.byref(n))
    return int(n.value)


def sites_per_pass(L, H, threads):
    """Sites a thread updates in the largest pass of a sweep (the kernel keeps
    their streams in registers and supports at most two)."""
    rows = threads // (L // 2)
    return -(-max(H // 2, 4) // rows)


def candidate_configs(code, L, R, exchange):
    """(strip height H, workgroup size) pairs whose grid (R * L / H workgroups)
    can be co-resident: H an even divisor of L, at most two sites per thread
    and pass, and at least one colour row per wave for 512 threads."""
    out = []
    for threads in WORKGROUP_SIZES:
        if L // 2 > threads:
            continue
        for H in range(4, L + 1, 2):
            if L % H or sites_per_pass(L, H, threads) > 2:
                continue
            if threads > 256 and (H // 2) * (L // 2) < threads:
                continue
            if R * (L // H) > max_workgroups(code, L, R, H, exchange, threads):
                continue
            out.append((H, threads))
    return out


def candidate_strip_rows(code, L, R, exchange):
    """Strip heights usable with 256-thread workgroups (or 512 when L > 512)."""
    return sorted({H for H, t in candidate_configs(code, L, R, exchange) if t == _min_threads(L)})


def _min_threads(L):
    return 256 if L // 2 <= 256 else 512


def default_config(code, L, R, exchange, target_workgroups=512):
    """Smallest strip height >= 4 (smallest workgroup) whose grid does not exceed `target_workgroups`."""
    cands = candidate_configs(code, L, R, exchange)
    if not cands:
        raise ValueError(f"no strip height fits L={L}, R={R} on this device")
    t0 = min(t for _, t in cands)
    base = [H for H, t in cands if t == t0]
    good = [H for H in base if R * (L // H) <= target_workgroups]
    return (min(good) if good else max(base)), t0


def default_strip_rows(code, L, R, exchange, target_workgroups=512):
    return default_config(code, L, R, exchange, target_workgroups)[0]


class Ensemble:
    """One model, one lattice size, all its replicas: state, RNG streams, PT
    state, proposal widths and acceptance counters on the device, with the
    host bookkeeping that launches advance.  ``BatchedRun`` runs one ensemble
    per launch; ``core.packed.PackedRun`` runs several in one launch."""

    def __init__(self, model, L, temperatures, seed=42, exchange_interval=10, pt_seed=None):
        if L < 4 or L % 2 != 0:
            raise ValueError("L must be even and >= 4")
        temperatures = np.asarray(temperatures, dtype=np.float64)
        if temperatures.ndim != 1 or len(temperatures) == 0:
            raise ValueError("Temperatures must be a nonempty 1D array")
        if not np.all(np.isfinite(temperatures) & (temperatures > 0)):
            raise ValueError("Temperatures must be positive and finite")
        self.code = model_code(model)
        if isinstance(exchange_interval, (bool, np.bool_)):
            raise ValueError("exchange_interval must be a nonnegative integer")
        exchange_interval = operator.index(exchange_interval)
        if exchange_interval < 0:
            raise ValueError("exchange_interval must be a nonnegative integer")
        self.model = model
        self.L = L
        self.R = len(temperatures)
        self.temperatures = temperatures.copy()
        self.seed = seed
        self.exchange_interval = exchange_interval
        betas = (1.0 / temperatures).astype(np.float32)
        if exchange_interval:
            if self.R < 2:
                raise ValueError("Parallel tempering requires at least two temperatures")
            if not np.all(np.diff(betas) < 0):
                raise ValueError("PT temperatures must have distinct, finite positive float32 betas")

        state = np.ascontiguousarray(model.initialize_state(L, self.R, seed))
        if state.shape != (self.R, L, L):
            raise ValueError("Invalid model state shape")
        self.state = to_device(state)
        self.rng_states = create_xoroshiro128p_states(self.R * L * L, seed=seed)
        self.betas = to_device(betas)
        self._temperature_betas = betas.copy()

        R = self.R
        self.exchange_seed = operator.index(seed + 1 if pt_seed is None else pt_seed)
        ident = np.arange(R, dtype=np.int32)
        self.walker_to_temperature = to_device(ident)
        self.temperature_to_walker = to_device(ident)
        self.energies = device_zeros(R, np.float64)
        self._pt_rng = create_xoroshiro128p_states(max(R - 1, 1), seed=self.exchange_seed)
        self._swap_counts = device_zeros((3, max(R - 1, 1)), np.int64)
        trips = np.zeros((5, R), dtype=np.int64)
        trips[0] = -1
        trips[2] = -1
        trips[0, 0], trips[0, -1] = 0, R - 1
        trips[2, [0, -1]] = 0
        self._trips = to_device(trips)
        self._parity = device_zeros(1, np.int32)
        self.sweeps = 0
        self.exchange_rounds = 0
        self._accepted = device_zeros(R, np.uint64)
        self._counted_sweeps = 0
        self.sigmas = None
        if self.code == MODEL_XY_GAUSS:
            self._sigmas_host = np.broadcast_to(np.asarray(model.sigma, dtype=np.float64), (R,)).copy()
            self.sigmas = to_device(self._sigmas_host.astype(np.float32))
        self._owner = None  # weak reference to the engine that runs this ensemble

    # ------------------------------------------------------- bookkeeping
    def _sample_plan(self, nsweeps, sample_interval, sample_first):
        """(first sampled sweep, number of samples) of a launch of `nsweeps`."""
        if sample_interval and sample_interval < 1:
            raise ValueError("sample_interval must be positive")
        if not sample_interval:
            return 0, 0
        sweep0 = self.sweeps
        if sample_first is None:
            sample_first = sweep0 + sample_interval
        elif sample_first <= sweep0:
            # samples are taken after sweeps of this launch only: the first
            # sweep of the series sample_first + k * interval past sweep0
            sample_first += (sweep0 - sample_first) // sample_interval * sample_interval + sample_interval
        last = sweep0 + nsweeps
        return sample_first, (0 if last < sample_first else (last - sample_first) // sample_interval + 1)

    def _advance(self, nsweeps):
        sweep0 = self.sweeps
        self.sweeps += nsweeps
        self._counted_sweeps += nsweeps
        if self.exchange_interval:
            self.exchange_rounds += (sweep0 + nsweeps) // self.exchange_interval - sweep0 // self.exchange_interval

    def _samples(self, partials, sw2t, n, first, interval):
        """Samples in temperature order from per-strip partials (n, R, S, C),
        strips added in order, and the walkers' temperatures (n, R)."""
        names = NAMES[self.code]
        if n == 0:
            return {name: np.zeros((0, self.R)) for name in names}, np.zeros(0, dtype=np.int64)
        walker = partials[:n].sum(axis=2)
        w2t = sw2t[:n]
        out = {}
        rows = np.arange(n)[:, None]
        for c, name in enumerate(names):
            values = np.empty((n, self.R))
            values[rows, w2t] = walker[:, :, c]
            out[name] = values
        return out, first + interval * np.arange(n, dtype=np.int64)

    @property
    def owner(self):
        """The packed engine that runs this ensemble, if any."""
        return self._owner() if self._owner is not None else None

    def synchronize(self):
        owner = self.owner
        if owner is not None:
            owner.synchronize()
        else:
            synchronize()

    # -------------------------------------------------- proposals, acceptance
    def acceptance_by_temperature(self):
        """Accepted fraction of site updates per ladder temperature since the last reset."""
        self.synchronize()
        acc = self._accepted.copy_to_host().astype(np.float64)
        attempts = float(self.L * self.L * self._counted_sweeps)
        return acc / attempts if attempts else np.full(self.R, np.nan)

    def reset_acceptance(self):
        self._accepted.fill_zero()
        self._counted_sweeps = 0

    def set_sigmas(self, sigmas):
        from models.xy.proposals import SIGMA_MAX, SIGMA_MIN
        if self.code != MODEL_XY_GAUSS:
            raise ValueError("proposal widths exist only for Gaussian XY proposals")
        sig = np.clip(np.broadcast_to(np.asarray(sigmas, dtype=np.float64), (self.R,)), SIGMA_MIN, SIGMA_MAX)
        self._sigmas_host = sig.copy()
        self.sigmas.copy_to_device(sig.astype(np.float32))

    def _adapt_step(self, target, rate):
        rates = self.acceptance_by_temperature()
        self.set_sigmas(self._sigmas_host * np.exp(rate * (rates - target)))
        self.reset_acceptance()
        return self._sigmas_host.copy(), rates

    # ---------------------------------------------------------- inspection
    def snapshot(self):
        """Host state in walker order."""
        return self.state.copy_to_host()

    def measure(self):
        """Observables of the current state through the drop-in measurement kernels,
        in temperature order."""
        return self.model.measure(self.state, temperature_indices=self.walker_to_temperature)

    def diagnostics(self):
        counts = self._swap_counts.copy_to_host()
        if self.R > 1 and np.any(counts[2]):
            raise FloatingPointError("Nonfinite energies encountered during replica exchange")
        trips = self._trips.copy_to_host()
        attempts, accepted = counts[0, :self.R - 1], counts[1, :self.R - 1]
        rates = np.divide(accepted, attempts, out=np.zeros(self.R - 1), where=attempts > 0)
        completed = int(trips[3].sum())
        return {
            "exchange_interval": self.exchange_interval,
            "exchange_seed": self.exchange_seed,
            "sweeps": self.sweeps,
            "exchange_rounds": self.exchange_rounds,
            "next_parity": int(self._parity.copy_to_host()[0]),
            "diagnostics_scope": "since_wrapper_creation_including_thermalization",
            "temperature_betas": self._temperature_betas.copy(),
            "attempts": attempts.copy(),
            "accepted": accepted.copy(),
            "acceptance_rates": rates,
            "walker_to_temperature": self.walker_to_temperature.copy_to_host(),
            "temperature_to_walker": self.temperature_to_walker.copy_to_host(),
            "walker_round_trips": trips[3].copy(),
            "walker_round_trip_total_sweeps": trips[4].copy(),
            "mean_round_trip_sweeps": float(trips[4].sum()) / completed if completed else None,
            "acceptance_by_temperature": self.acceptance_by_temperature(),
            "sigmas": None if self.sigmas is None else self._sigmas_host.copy(),
        }


class BatchedRun(Ensemble):
    """One ensemble in one persistent launch: every strip resident in LDS."""

    def __init__(self, model, L, temperatures, seed=42, exchange_interval=10, pt_seed=None,
                 strip_rows=None, stream=True, threads=None):
        super().__init__(model, L, temperatures, seed, exchange_interval, pt_seed)
        if strip_rows:
            self.H = int(strip_rows)
            self.threads = int(threads) if threads else _min_threads(L)
        else:
            self.H, t = default_config(self.code, L, self.R, self.exchange_interval)
            self.threads = int(threads) if threads else t
        self._alloc_workspace()
        self._stream = None
        if stream:
            p = ctypes.c_void_p()
            _lib.check(_lib.load().mc_stream_create(ctypes.byref(p)))
            self._stream = p

    def _alloc_workspace(self):
        hw, ew = ctypes.c_size_t(), ctypes.c_size_t()
        _lib.call("mc_batch_workspace_size", self.code, self.L, self.R, self.H, ctypes.byref(hw), ctypes.byref(ew))
        self._halo = device_array(max(hw.value, 1), np.uint64)
        self._energy_words = device_array(max(ew.value, 1), np.uint64)
        self._error = device_zeros(1, np.int32)  # the watchdog code
        # tags issued into the workspace since it was zeroed (0: zero it at
        # the next launch); the launches keep it
        self._clock = (ctypes.c_longlong * 1)()
        self._error_host = PinnedArray(1, np.int32)  # the watchdog code after each launch
        self._no_samples = (device_array((1, self.R, self.strips, CHANNELS), np.float64),
                            device_array((1, self.R), np.int32))
        limit = max_workgroups(self.code, self.L, self.R, self.H, self.exchange_interval, self.threads)
        if self.R * (self.L // self.H) > limit:
            raise ValueError(f"grid of {self.R * (self.L // self.H)} workgroups exceeds the "
                             f"{limit} that can be co-resident; choose a taller strip")

    @property
    def strips(self):
        return self.L // self.H

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing() or _lib.closing():
            return
        plan = getattr(self, "_plan", None)
        if plan is not None:
            _lib.load().mc_batch_plan_destroy(plan[1])
        s = getattr(self, "_stream", None)
        if s is not None:
            _lib.load().mc_stream_destroy(s)

    # ------------------------------------------------------------- running
    def launch(self, nsweeps, sample_interval=0, sample_first=None):
        """Queue `nsweeps` sweeps (asynchronous).  Returns a handle for collect()."""
        nsweeps = int(nsweeps)
        if nsweeps < 0:
            raise ValueError("nsweeps must be nonnegative")
        sample_first, n_samples = self._sample_plan(nsweeps, sample_interval, sample_first)
        sweep0 = self.sweeps
        S = self.strips
        if n_samples:
            partials = device_array((n_samples, self.R, S, CHANNELS), np.float64)
            sw2t = device_array((n_samples, self.R), np.int32)
        else:  # no samples: the kernel takes no writes, so one placeholder pair serves every launch
            partials, sw2t = self._no_samples
        _lib.check(_lib.load().mc_batch_plan_launch(self._launch_plan(), nsweeps, sweep0, int(sample_interval or 0),
                                                    int(sample_first), n_samples, as_pointer(partials),
                                                    as_pointer(sw2t)))
        self._advance(nsweeps)
        return (partials, sw2t, n_samples, sample_first, sample_interval)

    def _launch_plan(self):
        """The native launch plan (mc_batch_plan_create): the arguments that do not
        change between launches, checked once.  Made again when J, delta or one of
        the exchange interval or one of the arrays changes (the arrays keep their
        buffers for the run's life)."""
        arrays = (self.state, self.rng_states, self.betas, self.sigmas, self._accepted, self._parity, self.energies,
                  self.walker_to_temperature, self.temperature_to_walker, self._pt_rng, self._swap_counts,
                  self._trips, self._halo, self._energy_words, self._error)
        key = (float(self.model.J), float(getattr(self.model, "delta", 0.0)), self.exchange_interval,
               tuple(map(id, arrays)))
        cached = getattr(self, "_plan", None)
        if cached is not None and cached[0] == key:
            return cached[1]
        plan = ctypes.c_void_p()
        _lib.call("mc_batch_plan_create", self._stream, self.code, self.L, self.R, self.H, self.threads,
                  as_pointer(self.state), as_pointer(self.rng_states), as_pointer(self.betas), key[0], key[1],
                  as_pointer(self.sigmas), as_pointer(self._accepted), self.exchange_interval,
                  as_pointer(self._parity), as_pointer(self.energies), as_pointer(self.walker_to_temperature),
                  as_pointer(self.temperature_to_walker), as_pointer(self._pt_rng), as_pointer(self._swap_counts),
                  as_pointer(self._trips), as_pointer(self._halo), as_pointer(self._energy_words),
                  as_pointer(self._error), ctypes.cast(self._clock, ctypes.c_void_p), self._error_host.pointer,
                  ctypes.byref(plan))
        if cached is not None:
            _lib.load().mc_batch_plan_destroy(cached[1])
        self._plan = (key, plan)
        return plan

    def synchronize(self):
        if self._stream is not None:
            _lib.check(_lib.load().mc_stream_synchronize(self._stream))
        else:
            synchronize()
        err = int(self._error_host.host[0])
        if err:
            self._error_host.host[0] = 0
            self._clock[0] = 0  # the next launch zeroes the workspace
            raise RuntimeError(f"batched launch stopped by its watchdog (code {err}): "
                               "a workgroup waited too long for a neighbour or the barrier")

    def collect(self, handle):
        """Wait for a launch and return its samples in temperature order."""
        partials, sw2t, n, first, interval = handle
        self.synchronize()
        if n == 0:
            return self._samples(None, None, 0, first, interval)
        return self._samples(partials.copy_to_host(), sw2t.copy_to_host(), n, first, interval)

    def run(self, nsweeps, sample_interval=0, sample_first=None):
        return self.collect(self.launch(nsweeps, sample_interval, sample_first))

    def adapt_sigmas(self, sweeps, interval=50, target=0.5, rate=1.0):
        """Thermalise `sweeps` sweeps, moving each temperature's width toward `target`
        acceptance every `interval` sweeps (log sigma += rate (acceptance - target)).
        Returns [(sigmas, acceptance)] per step; widths stay fixed afterwards."""
        if self.code != MODEL_XY_GAUSS:
            raise ValueError("proposal widths exist only for Gaussian XY proposals")
        if interval < 1:
            raise ValueError("interval must be positive")
        history = []
        self.reset_acceptance()
        done = 0
        while done < sweeps:
            n = min(interval, sweeps - done)
            self.run(n)
            done += n
            history.append(self._adapt_step(target, rate))
        return history

    def diagnostics(self):
        d = super().diagnostics()
        d["strip_rows"] = self.H
        d["workgroup_threads"] = self.threads
        return d


def autotune(model, L, temperatures, seed=42, exchange_interval=10, sweeps=300, candidates=None,
             verbose=False):
    """Time short runs for each candidate (strip height, workgroup size) and return the fastest.

    Results are cached per (model type, L, R, exchange) for the process.
    """
    code = model_code(model)
    R = len(temperatures)
    key = (code, L, R, bool(exchange_interval))
    if key in _tuned and candidates is None:
        return _tuned[key]
    cands = candidates or candidate_configs(code, L, R, exchange_interval)
    best, best_t = None, None
    for H, threads in cands:
        run = BatchedRun(model, L, temperatures, seed=seed, exchange_interval=exchange_interval,
                         strip_rows=H, threads=threads)
        run.run(20)
        t0 = time.perf_counter()
        run.run(sweeps)
        dt = time.perf_counter() - t0
        if verbose:
            print(f"  H={H:4d} threads={threads:4d} strips={L // H:4d} {dt / sweeps * 1e6:9.2f} us/sweep",
                  file=sys.stderr, flush=True)
        if best_t is None or dt < best_t:
            best, best_t = (H, threads), dt
    _tuned[key] = best
    return best


def autotune_strip_rows(model, L, temperatures, seed=42, exchange_interval=10, sweeps=300, candidates=None,
                        verbose=False):
    """The fastest strip height among 256-thread configurations (see autotune)."""
    cands = candidates and [(H, _min_threads(L)) for H in candidates]
    return autotune(model, L, temperatures, seed, exchange_interval, sweeps, cands, verbose)[0]


def run_metadata(ens, sample_sweeps, thermalization_sweeps, measurement_sweeps, sample_interval, engine):
    """Run metadata in the reference's layout (core/simulation.py:80-91)."""
    return {
        "L": ens.L,
        "temperatures": ens.temperatures.copy(),
        "seed": ens.seed,
        "model_parameters": ens.model.parameters(),
        "thermalization_sweeps": thermalization_sweeps,
        "measurement_sweeps": measurement_sweeps,
        "sample_interval": sample_interval,
        "sample_sweeps": sample_sweeps,
        "engine": engine,
    }


def group_waves(shares):
    """First-fit grouping of launches into waves whose shares of the device
    (workgroups / co-resident limit) add up to at most one; a share above one
    gets a wave of its own (its launch then fails its own co-residency check)."""
    waves, load = [], []
    for i, share in enumerate(shares):
        for w, used in enumerate(load):
            if used + share <= 1.0:
                waves[w].append(i)
                load[w] += share
                break
        else:
            waves.append([i])
            load.append(share)
    return waves


class BatchedSimulation:
    """Thermalise and sample several BatchedRuns, concurrently (one stream each)
    as far as their grids fit on the device together (``waves``).

    Mirrors core/simulation.py: ``sample_sweeps = thermalization + step`` for
    every ``step % sample_interval == 0``; returns ``{L: {"samples", "metadata"}}``.
    """

    def __init__(self, runs, thermalization_sweeps=1000, measurement_sweeps=4000, sample_interval=10):
        if thermalization_sweeps < 0:
            raise ValueError("Thermalization sweeps cannot be negative")
        if measurement_sweeps < 1:
            raise ValueError("Measurement sweeps must be positive")
        if sample_interval < 1 or sample_interval > measurement_sweeps:
            raise ValueError("Invalid sampling interval")
        self.runs = list(runs)
        self.thermalization_sweeps = thermalization_sweeps
        self.measurement_sweeps = measurement_sweeps
        self.sample_interval = sample_interval

    def waves(self):
        """Indices of the runs grouped into waves that can run at the same time.

        Each run is a persistent launch whose workgroups wait for each other,
        so runs in flight together must fit on the device together: a wave
        holds runs whose grids, as fractions of their own co-residency limits
        (mc_batch_max_workgroups), add up to at most one.  Waves run one
        after another; runs of a wave run concurrently on their own streams.
        """
        shares = []
        for br in self.runs:
            limit = max_workgroups(br.code, br.L, br.R, br.H, br.exchange_interval, br.threads)
            shares.append(br.R * br.strips / limit)
        return group_waves(shares)

    def run(self):
        total = self.thermalization_sweeps + self.measurement_sweeps
        results = [None] * len(self.runs)
        for wave in self.waves():
            handles = []
            for i in wave:
                br = self.runs[i]
                first = br.sweeps + self.thermalization_sweeps + self.sample_interval
                handles.append((i, br.launch(total, self.sample_interval, first)))
            for i, h in handles:
                results[i] = self.runs[i].collect(h)
        out = {}
        for br, (samples, sweeps) in zip(self.runs, results):
            metadata = run_metadata(br, sweeps - (br.sweeps - total), self.thermalization_sweeps,
                                    self.measurement_sweeps, self.sample_interval, "batched")
            if br.exchange_interval:
                metadata["parallel_tempering"] = br.diagnostics()
            key = br.L
            while key in out:  # several models at one L: keep all
                key = (key, len(out))
            out[key] = {"samples": samples, "metadata": metadata}
        return out
