"""Kawamura (PRL 82, 964, 1999) campaign with parallel tempering
(main/kawamura.py of the reference).  Run from gpu-monte-carlo_amd/; each
lattice size is saved on its own as data/kawamura/<campaign>/L=<size>.npz."""

from pathlib import Path
from time import perf_counter

import numpy as np

from core.parallel_tempering import ParallelTempering
from core.runtime import Runtime
from core.simulation import Simulation
from main._campaign import preflight, save_atomically
from models.kawamura.model import KawamuraModel
from utils.temperature_ladder import make_temperature_ladder

# physical / model parameters
J = 1.0
delta = 1.5                 # phase proposal half-width (0 < delta <= pi)
sizes = [64, 80, 96]        # even L >= 4
dtype = np.float32          # np.float64 for the float64 mode

# sampling
thermalization_sweeps = 500_000
measurement_sweeps = 1_000_000
sample_interval = 10
exchange_interval = 10

# GPU and parallel tempering
threads = 16
seed_base = 42
pt_seed_base = 43
tracked_walkers = "all"     # "all", None, or a list of walker IDs
history_stride = 10         # exchange rounds

# lower transition T/J ~ 0.692, chiral transition T/J ~ 0.753
temperature_ladder = {"T_min": 0.65 * J, "T_max": 0.80 * J, "n_T": 100, "Tc": None}

campaign = "kawamura_reproduction_01"
data_root = Path("data/kawamura")


def output_path(L):
    return data_root / campaign / f"L={L}.npz"


def main():
    if not np.isfinite(J) or J <= 0:
        raise ValueError("J must be positive and finite")
    if not np.isfinite(delta) or not 0 < delta <= np.pi:
        raise ValueError("Require 0 < delta <= pi")
    if not sizes or any(L < 4 or L % 2 for L in sizes):
        raise ValueError("All lattice sizes must be even and >= 4")
    temperatures = make_temperature_ladder(**temperature_ladder)
    walker_ids = range(len(temperatures)) if tracked_walkers == "all" else tracked_walkers
    preflight((output_path(L) for L in sizes), "Duplicate simulation output paths (check sizes)")

    print(f"\nKawamura simulation: {campaign}")
    print(f"J={J:g}, delta={delta:g}, sizes {sizes}, {len(temperatures)} replicas")
    saved = []
    for L in sizes:
        print(f"\nL={L}")
        runtime = Runtime(model=KawamuraModel(J=J, delta=delta, dtype=dtype), L=L,
                          temperatures=temperatures, threads=threads, seed=seed_base + L)
        pt = ParallelTempering(runtime, exchange_interval=exchange_interval, seed=pt_seed_base + L,
                               tracked_walkers=walker_ids, history_stride=history_stride)
        simulation = Simulation(runtime=pt, thermalization_sweeps=thermalization_sweeps,
                                measurement_sweeps=measurement_sweeps, sample_interval=sample_interval)
        start = perf_counter()
        run = simulation.run()
        seconds = perf_counter() - start
        metadata = run["metadata"]
        metadata["timing"] = {"simulation_seconds": seconds}
        metadata["campaign"] = {"id": campaign, "sizes": list(sizes)}
        metadata["sampler"] = "metropolis"
        metadata["threads"] = threads
        metadata["temperature_ladder"] = dict(temperature_ladder)
        metadata["sampler_parameters"] = {"delta": delta}
        start = perf_counter()
        saved.append(save_atomically(output_path(L), L, run))
        print(f"  Simulation: {seconds / 60:.2f} min")
        print(f"  Saving:     {perf_counter() - start:.1f} s")
        print(f"  Saved:      {saved[-1]}")
    return saved


if __name__ == "__main__":
    main()
