"""Packed persistent engine: the runs of a whole finite-size-scaling set --
several lattice sizes and models, each with its own temperature ladder and
parallel tempering -- in one launch (``mc_pack_run``), optionally captured in
a HIP graph and replayed.

Every ensemble (``core.batched.Ensemble``) keeps the reference's schedule:
sweeps (core/runtime.py:87-99), an exchange every ``exchange_interval``
sweeps (core/parallel_tempering.py:245-269) and samples every
``sample_interval`` sweeps (core/simulation.py:36-72).  Trajectories, swaps
and samples are those of the same ensemble run alone by ``BatchedRun``: the
site updates are the same operations on the same streams, and energies are
integer sums, independent of how sites are spread over workgroups.

Planning (``plan_pack``) is host logic: per ensemble a strip height H and a
replica block K (a task is one strip position of K replicas, about
``task_sites`` sites), then a fixed number of tasks per workgroup chosen so
that every workgroup of the launch is co-resident.  ``autotune_pack`` times
candidate (workgroup size, task size) pairs on scratch copies of the
ensembles.
"""

import ctypes
import sys
import math
import time
import weakref

import numpy as np

from core.batched import CHANNELS, MODEL_ISING, MODEL_XY_GAUSS, Ensemble, run_metadata
from mcgpu import _lib, as_pointer, device_array

WORKGROUP_SIZES = (256, 512)
TASK_SITES = (1024, 2048, 4096, 8192)
DEFAULT_CONFIG = (256, 2048)


# ------------------------------------------------------------------ planning

def group_shapes(L, R, threads):
    """Every valid (H, K) of one ensemble: H an even divisor of L (>= 4), K a
    power of two dividing R with K strips fitting one pass, K rows of L sites
    fitting one row map (K * L <= NT) and NT / K >= 16 threads per replica
    for the observables."""
    if L > threads or L % 2 or L < 4:
        return []
    rows = threads // (L // 2)
    out = []
    for H in range(4, L + 1, 2):
        if L % H:
            continue
        K = 1
        while K <= threads // 16 and K <= rows and K * L <= threads:
            if R % K == 0:
                out.append((H, K))
            K *= 2
    return out


def choose_shape(L, R, threads, task_sites, lds_of=None, lds_limit=159 * 1024):
    """(H, K) whose task (K * H * L sites) is closest to `task_sites` in ratio;
    ties go to taller strips (less halo and ghost traffic), then fewer replicas."""
    best, key = None, None
    for H, K in group_shapes(L, R, threads):
        if lds_of is not None and lds_of(H, K) > lds_limit:
            continue
        k = (abs(math.log(K * H * L / task_sites)), -H, K)
        if key is None or k < key:
            best, key = (H, K), k
    if best is None:
        raise ValueError(f"no strip shape fits L={L}, R={R} with {threads}-thread workgroups")
    return best


def split_tasks(ntasks, per_workgroup):
    """Consecutive task ranges (first, count) over ceil(ntasks / per_workgroup)
    workgroups, counts differing by at most one."""
    nw = -(-ntasks // per_workgroup)
    base, extra = divmod(ntasks, nw)
    out, t = [], 0
    for i in range(nw):
        n = base + (1 if i < extra else 0)
        out.append((t, n))
        t += n
    return out


def plan_pack(groups, threads, task_sites, capacity_of):
    """Plan a packed launch.

    groups: [(code, L, R, lds_of)] with lds_of(H, K) -> bytes.
    capacity_of(lds_bytes) -> co-resident workgroups of that LDS size.
    Returns {"shapes": [(H, K)], "work": [(group, first task, count)],
    "tasks_per_workgroup": T, "lds": bytes}.
    """
    shapes = [choose_shape(L, R, threads, task_sites, lds_of) for _, L, R, lds_of in groups]
    ntasks = [(R // K) * (L // H) for (_, L, R, _), (H, K) in zip(groups, shapes)]
    lds = max(lds_of(H, K) for (_, _, _, lds_of), (H, K) in zip(groups, shapes))
    capacity = capacity_of(lds)
    if capacity <= 0:
        raise ValueError(f"a workgroup of {lds} bytes of LDS cannot run")
    T = 1
    while sum(-(-n // T) for n in ntasks) > capacity:
        T += 1
    work = []
    for g, n in enumerate(ntasks):
        work.extend((g, first, count) for first, count in split_tasks(n, T))
    return {"shapes": shapes, "work": work, "tasks_per_workgroup": T, "lds": lds}


def _lds_bytes(code, L, R, H, K):
    n = ctypes.c_size_t()
    _lib.call("mc_pack_lds_bytes", code, L, R, H, K, ctypes.byref(n))
    return int(n.value)


def _capacity(threads, lds):
    n = ctypes.c_int()
    _lib.call("mc_pack_occupancy", int(threads), ctypes.c_size_t(int(lds)), ctypes.byref(n))
    return int(n.value)


# ------------------------------------------------------------------ engine

class PackedRun:
    """Several ensembles advanced together by one launch per call (or one
    graph replay).  All ensembles must have run the same number of sweeps."""

    def __init__(self, ensembles, threads=None, task_sites=None):
        self.ensembles = list(ensembles)
        if not self.ensembles:
            raise ValueError("need at least one ensemble")
        for e in self.ensembles:
            if not isinstance(e, Ensemble):
                raise TypeError("PackedRun packs core.batched.Ensemble objects")
            if e.owner is not None:
                raise ValueError("an ensemble can belong to one engine only")
        sweeps = {e.sweeps for e in self.ensembles}
        if len(sweeps) != 1:
            raise ValueError("packed ensembles must have run the same number of sweeps")
        if threads is None or task_sites is None:
            t, s = autotune_pack([(e.model, e.L, e.temperatures, e.exchange_interval) for e in self.ensembles])
            threads = threads or t
            task_sites = task_sites or s
        self.threads, self.task_sites = int(threads), int(task_sites)
        self._pack = None
        self._stream = None
        groups = [(e.code, e.L, e.R, (lambda H, K, e=e: _lds_bytes(e.code, e.L, e.R, H, K)))
                  for e in self.ensembles]
        self.plan = plan_pack(groups, self.threads, self.task_sites, lambda lds: _capacity(self.threads, lds))

        lib = _lib.load()
        p = ctypes.c_void_p()
        _lib.check(lib.mc_pack_create(self.threads, ctypes.byref(p)))
        self._pack = p
        for e, (H, K) in zip(self.ensembles, self.plan["shapes"]):
            idx = ctypes.c_int()
            _lib.call("mc_pack_add_group", self._pack, e.code, e.L, e.R, H, K, e.exchange_interval,
                      float(e.model.J), float(getattr(e.model, "delta", 0.0)), as_pointer(e.state),
                      as_pointer(e.rng_states), as_pointer(e.betas), as_pointer(e.sigmas), as_pointer(e._accepted),
                      as_pointer(e._parity), as_pointer(e.energies), as_pointer(e.walker_to_temperature),
                      as_pointer(e.temperature_to_walker), as_pointer(e._pt_rng), as_pointer(e._swap_counts),
                      as_pointer(e._trips), ctypes.byref(idx))
            e.H, e.K = H, K
        work = np.asarray(self.plan["work"], dtype=np.int32).reshape(-1)
        _lib.call("mc_pack_set_work", self._pack, work.ctypes.data_as(ctypes.c_void_p), len(self.plan["work"]))
        _lib.call("mc_pack_set_sweeps", self._pack, int(sweeps.pop()))
        s = ctypes.c_void_p()
        _lib.check(lib.mc_stream_create(ctypes.byref(s)))
        self._stream = s
        for e in self.ensembles:
            e._owner = weakref.ref(self)

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing() or _lib.closing() or _lib._lib is None:
            return
        lib = _lib.load()
        if getattr(self, "_pack", None) is not None:
            lib.mc_pack_destroy(self._pack)
            self._pack = None
        if getattr(self, "_stream", None) is not None:
            lib.mc_stream_destroy(self._stream)
            self._stream = None

    @property
    def workgroups(self):
        return len(self.plan["work"])

    def describe(self):
        """Per ensemble: model code, L, R, strip rows, replicas per task, tasks, workgroups."""
        counts = {}
        for g, _, _ in self.plan["work"]:
            counts[g] = counts.get(g, 0) + 1
        return [{"model": e.code, "L": e.L, "R": e.R, "strip_rows": H, "replicas_per_task": K,
                 "tasks": (e.R // K) * (e.L // H), "workgroups": counts[i]}
                for i, (e, (H, K)) in enumerate(zip(self.ensembles, self.plan["shapes"]))]

    # ------------------------------------------------------------- running
    def _buffers(self, n_samples):
        parts, w2ts = [], []
        for e in self.ensembles:
            parts.append(device_array((max(n_samples, 1), e.R, e.L // e.H, CHANNELS), np.float64))
            w2ts.append(device_array((max(n_samples, 1), e.R), np.int32))
        pp = (ctypes.c_void_p * len(parts))(*[as_pointer(x) for x in parts])
        pw = (ctypes.c_void_p * len(w2ts))(*[as_pointer(x) for x in w2ts])
        return parts, w2ts, pp, pw

    def launch(self, nsweeps, sample_interval=0, sample_first=None):
        """Queue `nsweeps` sweeps of every ensemble (asynchronous); returns a handle for collect()."""
        nsweeps = int(nsweeps)
        if nsweeps < 0:
            raise ValueError("nsweeps must be nonnegative")
        first, n = self.ensembles[0]._sample_plan(nsweeps, sample_interval, sample_first)
        parts, w2ts, pp, pw = self._buffers(n)
        _lib.call("mc_pack_run", self._pack, self._stream, nsweeps, int(sample_interval or 0), int(first), n, pp, pw)
        for e in self.ensembles:
            e._advance(nsweeps)
        return (parts, w2ts, n, first, sample_interval)

    def synchronize(self):
        _lib.check(_lib.load().mc_stream_synchronize(self._stream))
        err = ctypes.c_int()
        _lib.call("mc_pack_error", self._pack, ctypes.byref(err))
        if err.value:
            raise RuntimeError(f"packed launch stopped by its watchdog (code {err.value}): "
                               "a workgroup waited too long for a neighbour or an exchange")

    def collect(self, handle, n=None):
        """Wait and return [(samples, sample_sweeps)] per ensemble, in temperature order."""
        parts, w2ts, n_all, first, interval = handle
        n = n_all if n is None else n
        self.synchronize()
        out = []
        for e, p, w in zip(self.ensembles, parts, w2ts):
            if n == 0:
                out.append(e._samples(None, None, 0, first, interval))
            else:
                out.append(e._samples(p.copy_to_host(), w.copy_to_host(), n, first, interval))
        return out

    def run(self, nsweeps, sample_interval=0, sample_first=None):
        return self.collect(self.launch(nsweeps, sample_interval, sample_first))

    def capture(self, nsweeps, replays, sample_interval=0, sample_first=None):
        """A HIP graph of one launch of `nsweeps` sweeps, to be replayed up to
        `replays` times; samples of all replays land in one buffer."""
        return PackedGraph(self, int(nsweeps), int(replays), sample_interval, sample_first)

    def adapt_sigmas(self, sweeps, interval=50, target=0.5, rate=1.0):
        """Adapt the Gaussian widths of every Gaussian-XY ensemble (see BatchedRun.adapt_sigmas)."""
        gauss = [e for e in self.ensembles if e.code == MODEL_XY_GAUSS]
        if not gauss:
            raise ValueError("no ensemble has Gaussian proposals")
        if interval < 1:
            raise ValueError("interval must be positive")
        history = {id(e): [] for e in gauss}
        for e in gauss:
            e.reset_acceptance()
        done = 0
        while done < sweeps:
            n = min(interval, sweeps - done)
            self.run(n)
            done += n
            for e in gauss:
                history[id(e)].append(e._adapt_step(target, rate))
        return [history[id(e)] for e in gauss]

    def diagnostics(self, index):
        e = self.ensembles[index]
        d = e.diagnostics()
        d["strip_rows"], d["replicas_per_task"] = e.H, e.K
        d["workgroup_threads"] = self.threads
        d["packed_workgroups"] = self.workgroups
        return d


class PackedGraph:
    """A captured prologue + packed launch; each replay continues every chain
    by `nsweeps` sweeps.  The sample buffers cover all `replays` replays."""

    def __init__(self, run, nsweeps, replays, sample_interval=0, sample_first=None):
        if nsweeps < 1 or replays < 1:
            raise ValueError("nsweeps and replays must be positive")
        self.run, self.nsweeps, self.replays = run, nsweeps, replays
        self.done = 0
        e0 = run.ensembles[0]
        self.first, self.n_max = e0._sample_plan(nsweeps * replays, sample_interval, sample_first)
        self.interval = sample_interval
        self._sweep0 = e0.sweeps
        self.parts, self.w2ts, pp, pw = run._buffers(self.n_max)
        g = ctypes.c_void_p()
        _lib.call("mc_pack_capture", run._pack, run._stream, nsweeps, int(sample_interval or 0), int(self.first),
                  self.n_max, pp, pw, ctypes.byref(g))
        self._graph = g

    def __del__(self, _finalizing=sys.is_finalizing):
        if getattr(self, "_graph", None) is not None and not _finalizing() and not _lib.closing():
            _lib.load().mc_graph_destroy(self._graph)
            self._graph = None

    def replay(self):
        if self.done >= self.replays:
            raise ValueError("all replays of this graph are used (its sample buffers are full)")
        if any(e.sweeps != self._sweep0 + self.done * self.nsweeps for e in self.run.ensembles):
            raise ValueError("the ensembles were advanced outside this graph")
        _lib.call("mc_graph_launch", self._graph, self.run._stream)
        self.done += 1
        for e in self.run.ensembles:
            e._advance(self.nsweeps)

    def collect(self):
        """Samples of the replays so far: [(samples, sample_sweeps)] per ensemble."""
        last = self._sweep0 + self.done * self.nsweeps
        n = 0 if not self.interval or last < self.first else min(self.n_max, (last - self.first) // self.interval + 1)
        return self.run.collect((self.parts, self.w2ts, self.n_max, self.first, self.interval), n=n)


def autotune_pack(specs, sweeps=60, candidates=None, verbose=False, seed=1):
    """Fastest (workgroup size, task sites) for ensembles [(model, L, temperatures,
    exchange_interval)], timed on scratch ensembles of the same shapes."""
    cands = candidates or [(t, s) for t in WORKGROUP_SIZES for s in TASK_SITES]
    scratch = [Ensemble(m, L, T, seed=seed, exchange_interval=x) for m, L, T, x in specs]
    best, best_t = None, None
    for threads, sites in cands:
        try:
            pr = PackedRun(scratch, threads=threads, task_sites=sites)
        except ValueError:
            continue
        pr.run(5)
        t0 = time.perf_counter()
        pr.run(sweeps)
        dt = (time.perf_counter() - t0) / sweeps
        if verbose:
            print(f"  threads={threads:4d} task_sites={sites:6d} workgroups={pr.workgroups:5d} "
                  f"{dt * 1e6:9.2f} us/sweep", file=sys.stderr, flush=True)
        if best_t is None or dt < best_t:
            best, best_t = (threads, sites), dt
        del pr  # frees the pack; the scratch ensembles are free to join the next one
    if best is None:
        raise ValueError("no packed configuration fits these ensembles")
    return best


class PackedSimulation:
    """Thermalise and sample every ensemble of a PackedRun (core/simulation.py:
    ``sample_sweeps = thermalization + step``).  With ``chunk`` the sweeps run
    as replays of one captured graph of `chunk` sweeps; otherwise as one launch.
    Returns ``{L: {"samples", "metadata"}}`` (``(L, i)`` keys when sizes repeat)."""

    def __init__(self, packed, thermalization_sweeps=1000, measurement_sweeps=4000, sample_interval=10, chunk=None):
        if thermalization_sweeps < 0:
            raise ValueError("Thermalization sweeps cannot be negative")
        if measurement_sweeps < 1:
            raise ValueError("Measurement sweeps must be positive")
        if sample_interval < 1 or sample_interval > measurement_sweeps:
            raise ValueError("Invalid sampling interval")
        total = thermalization_sweeps + measurement_sweeps
        if chunk is not None and (chunk < 1 or total % chunk):
            raise ValueError("chunk must divide thermalization + measurement sweeps")
        self.packed = packed
        self.thermalization_sweeps = thermalization_sweeps
        self.measurement_sweeps = measurement_sweeps
        self.sample_interval = sample_interval
        self.chunk = chunk

    def run(self):
        total = self.thermalization_sweeps + self.measurement_sweeps
        pr = self.packed
        start = pr.ensembles[0].sweeps
        first = start + self.thermalization_sweeps + self.sample_interval
        if self.chunk:
            graph = pr.capture(self.chunk, total // self.chunk, self.sample_interval, first)
            for _ in range(total // self.chunk):
                graph.replay()
            results = graph.collect()
        else:
            results = pr.run(total, self.sample_interval, first)
        out = {}
        for i, (e, (samples, sweeps)) in enumerate(zip(pr.ensembles, results)):
            metadata = run_metadata(e, sweeps - start, self.thermalization_sweeps, self.measurement_sweeps,
                                    self.sample_interval, "packed")
            if e.exchange_interval:
                metadata["parallel_tempering"] = pr.diagnostics(i)
            key = e.L
            while key in out:
                key = (key, len(out))
            out[key] = {"samples": samples, "metadata": metadata}
        return out


__all__ = ["PackedRun", "PackedGraph", "PackedSimulation", "plan_pack", "choose_shape", "group_shapes",
           "split_tasks", "autotune_pack", "MODEL_ISING"]
