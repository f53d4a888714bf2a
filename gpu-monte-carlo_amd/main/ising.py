"""2D Ising campaign (main/ising.py of the reference): a Tc-focused ladder,
parallel tempering every 10 sweeps, one Runtime per lattice size, results in
data/ising_run.npz.  ``main(batched=True)`` runs every size in one concurrent
batch on the persistent engine instead (same samples, bit for bit)."""

from pathlib import Path

import numpy as np

from core.parallel_tempering import ParallelTempering
from core.runtime import Runtime
from core.simulation import Simulation
from models.ising.model import IsingModel
from utils.data_io import save_runs
from utils.temperature_ladder import make_temperature_ladder

sizes = [48, 64, 80]
J = 1.0
thermalization_sweeps = 100000
measurement_sweeps = 500000
sample_interval = 10
exchange_interval = 10
data_path = Path("data/ising_run.npz")


def ladder():
    Tc = 2 * J / np.log(1 + np.sqrt(2))
    return make_temperature_ladder(T_min=2.0, T_max=2.5, n_T=100, Tc=Tc, tc_window=0.15,
                                   tc_fraction=0.60)


def main(batched=False):
    temperatures = ladder()
    runs = {}
    if batched:
        from core.batched import BatchedRun, BatchedSimulation
        batch = [BatchedRun(IsingModel(J=J), L, temperatures, seed=42 + L,
                            exchange_interval=exchange_interval, pt_seed=43) for L in sizes]
        runs = BatchedSimulation(batch, thermalization_sweeps, measurement_sweeps, sample_interval).run()
    else:
        for L in sizes:
            print(f"Running Ising model: L = {L}")
            runtime = Runtime(model=IsingModel(J=J), L=L, temperatures=temperatures, threads=16,
                              seed=42 + L)
            pt = ParallelTempering(runtime, exchange_interval=exchange_interval, seed=43)
            runs[L] = Simulation(runtime=pt, thermalization_sweeps=thermalization_sweeps,
                                 measurement_sweeps=measurement_sweeps,
                                 sample_interval=sample_interval).run()
    save_runs(data_path, runs)
    print(f"Saved simulation to {data_path}")
    return data_path


if __name__ == "__main__":
    main()
