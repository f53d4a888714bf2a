"""Thermalise-then-sample driver.

Drop-in for core/simulation.py of the reference (constructor checks :10-21,
schedule :36-72, result layout :74-96): after ``thermalization_sweeps``
sweeps, a sample after every ``sample_interval``-th of the
``measurement_sweeps`` sweeps, recorded at global sweep
``thermalization_sweeps + step``; returns ``{"samples", "metadata"}`` with a
(samples, R) array per observable and the reference's metadata (plus the
exchange diagnostics when the runtime is a ParallelTempering).

How it is built here: the run is a list of segments -- sweeps, then a
sample -- handed to the runtime's ``advance(n)`` when it has one (a native
chain queues a whole segment in one call) and to ``sweep()`` one call at a
time otherwise.  Samples go to a (samples, C, R) float64 device buffer
through ``measure_into`` and come back in one copy at the end; a runtime
without a device measurement, or whose ``measure`` is replaced, is sampled
through ``measure()`` with the reference's checks.
"""

import numpy as np

from core.parallel_tempering import ParallelTempering
from mcgpu import device_array


def _stepper(runtime):
    """n -> run n sweeps; advance() unless sweep() is replaced on the instance."""
    advance = getattr(runtime, "advance", None)
    if advance is None or "sweep" in vars(runtime):
        def step(n):
            for _ in range(n):
                runtime.sweep()
        return step
    return advance


class _HostSamples:
    """Samples from runtime.measure(), checked as the reference checks them."""

    def __init__(self, runtime):
        self.runtime = runtime
        self.columns = None

    def take(self):
        values = self.runtime.measure()
        if not values:
            raise ValueError("Model returned no measurements")
        if self.columns is None:
            self.columns = {name: [] for name in values}
        if values.keys() != self.columns.keys():
            raise ValueError("Measurement channels changed")
        for name, v in values.items():
            v = np.asarray(v)
            if v.ndim < 1 or v.shape[0] != self.runtime.R:
                raise ValueError(f"Invalid replica dimension for {name}")
            self.columns[name].append(v.copy())

    def result(self):
        return {name: np.stack(rows, axis=0) for name, rows in (self.columns or {}).items()}


class _DeviceSamples:
    """Samples written by the measurement kernels into rows of one device buffer."""

    def __init__(self, runtime, count, channels):
        self.runtime = runtime
        self.rows = device_array((count, channels, runtime.R), dtype=np.float64)
        self.taken = 0
        self.names = None

    def take(self):
        self.names = self.runtime.measure_into(self.rows[self.taken])
        self.taken += 1

    def result(self):
        host = self.rows[:self.taken].copy_to_host()
        return {name: np.ascontiguousarray(host[:, c, :]) for c, name in enumerate(self.names)}


def _recorder(runtime, count):
    """Device rows when the runtime measures into them with the built-in
    measurement, else the host path."""
    channels = getattr(runtime, "sample_channels", None)
    names = None if channels is None or "measure" in vars(runtime) else channels()
    if names is None:
        return _HostSamples(runtime)
    return _DeviceSamples(runtime, count, len(names))


class Simulation:

    def __init__(self, runtime, thermalization_sweeps=1000, measurement_sweeps=4000, sample_interval=10):
        if thermalization_sweeps < 0:
            raise ValueError("Thermalization sweeps cannot be negative")
        if measurement_sweeps < 1:
            raise ValueError("Measurement sweeps must be positive")
        if sample_interval < 1 or sample_interval > measurement_sweeps:
            raise ValueError("Invalid sampling interval")
        self.runtime = runtime
        self.thermalization_sweeps = thermalization_sweeps
        self.measurement_sweeps = measurement_sweeps
        self.sample_interval = sample_interval

    def plan(self):
        """(sweeps before the first sample, sweeps between samples, number of
        samples, sweeps after the last sample)."""
        count = self.measurement_sweeps // self.sample_interval
        tail = self.measurement_sweeps - count * self.sample_interval
        return self.thermalization_sweeps, self.sample_interval, count, tail

    def run(self):
        rt = self.runtime
        lead, gap, count, tail = self.plan()
        step = _stepper(rt)
        samples = _recorder(rt, count)
        step(lead)
        for _ in range(count):
            step(gap)
            samples.take()
        step(tail)

        metadata = {
            "L": rt.L,
            "temperatures": rt.temperatures.copy(),
            "seed": rt.seed,
            "model_parameters": rt.model.parameters(),
            "thermalization_sweeps": self.thermalization_sweeps,
            "measurement_sweeps": self.measurement_sweeps,
            "sample_interval": self.sample_interval,
            "sample_sweeps": lead + gap * np.arange(1, count + 1),
        }
        if isinstance(rt, ParallelTempering):
            metadata["parallel_tempering"] = rt.diagnostics()
        return {"samples": samples.result(), "metadata": metadata}
