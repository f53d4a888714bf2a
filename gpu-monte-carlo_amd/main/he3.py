"""He3 thin-film campaign (main/he3.py of the reference): pressure and dipole
scans with the conditional-Gaussian or Cartesian sampler and parallel
tempering; each size saved atomically under data/he3/<campaign>/."""

from pathlib import Path
from time import perf_counter

from core.parallel_tempering import ParallelTempering
from core.runtime import Runtime
from core.simulation import Simulation
from main._campaign import preflight, save_atomically
from models.he3.model import He3Model
from models.he3.sampling import He3GaussianRuntime
from utils.temperature_ladder import make_temperature_ladder

# physical parameters
pressures = [20.0]
g_d_values = [0.001]
sizes = [64, 80, 96]
K = 1.0
betas = (-0.5, 1, 1, 1, -1)
delta = 0.15
reduced_temperature = 1.0

# sampling
sampler = "gaussian"  # "gaussian" or "cartesian"
gaussian_rho = -0.9
gaussian_proposal_mass = 0.0
thermalization_sweeps = 500_000
measurement_sweeps = 1_000_000
sample_interval = 10
exchange_interval = 10

# GPU and parallel tempering
threads = 16
seed_base = 42
pt_seed_base = 43
tracked_walkers = "all"
history_stride = 10

temperature_ladder = {"T_min": 0.365, "T_max": 0.415, "n_T": 100, "Tc": 0.385,
                      "tc_window": 0.01, "tc_fraction": 0.35}

campaign = "soc_scan_01"
data_root = Path("data/he3")


def output_path(pressure, g_d, L):
    return data_root / campaign / f"P={pressure:g}_gd={g_d:g}_{sampler}" / f"L={L}.npz"


def main():
    temperatures = make_temperature_ladder(**temperature_ladder)
    if sampler == "gaussian":
        runtime_class = He3GaussianRuntime
        runtime_options = {"rho": gaussian_rho, "proposal_mass": gaussian_proposal_mass}
    elif sampler == "cartesian":
        runtime_class, runtime_options = Runtime, {}
    else:
        raise ValueError(f"Unknown sampler: {sampler}")
    walker_ids = range(len(temperatures)) if tracked_walkers == "all" else tracked_walkers
    preflight(output_path(p, g, L) for p in pressures for g in g_d_values for L in sizes)

    saved = []
    print(f"\nHe3 Simulation: {campaign}")
    for pressure in pressures:
        for g_d in g_d_values:
            print(f"\nP={pressure:g} bar, g_d={g_d:g}, sampler={sampler}")
            for L in sizes:
                print(f"\nL={L}")
                model = He3Model(K=K, betas=betas, delta=delta, pressure=pressure,
                                 reduced_temperature=reduced_temperature, g_d=g_d)
                runtime = runtime_class(model=model, L=L, temperatures=temperatures, threads=threads,
                                        seed=seed_base + L, **runtime_options)
                pt = ParallelTempering(runtime, exchange_interval=exchange_interval,
                                       seed=pt_seed_base + L, tracked_walkers=walker_ids,
                                       history_stride=history_stride)
                simulation = Simulation(runtime=pt, thermalization_sweeps=thermalization_sweeps,
                                        measurement_sweeps=measurement_sweeps,
                                        sample_interval=sample_interval)
                start = perf_counter()
                run = simulation.run()
                seconds = perf_counter() - start
                metadata = run["metadata"]
                metadata["timing"] = {"simulation_seconds": seconds}
                metadata["campaign"] = {"id": campaign, "sizes": list(sizes)}
                metadata["sampler"] = sampler
                metadata["threads"] = threads
                metadata["temperature_ladder"] = dict(temperature_ladder)
                metadata["sampler_parameters"] = (dict(runtime_options) if sampler == "gaussian"
                                                  else {"delta": delta})
                if sampler == "gaussian":
                    metadata["sampling"] = runtime.sampling_diagnostics()
                start = perf_counter()
                saved.append(save_atomically(output_path(pressure, g_d, L), L, run))
                print(f"  Simulation: {seconds / 60:.2f} min")
                print(f"  Saving:     {perf_counter() - start:.1f} s")
                print(f"  Saved:      {saved[-1]}")
    return saved


if __name__ == "__main__":
    main()
