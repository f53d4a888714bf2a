"""2D XY campaign (main/xy.py of the reference): Tc-focused ladder around the
BKT temperature, parallel tempering every 10 sweeps, results in
data/xy_run.npz.  ``main(batched=True)`` uses the batched FP32 engine."""

from pathlib import Path

from core.parallel_tempering import ParallelTempering
from core.runtime import Runtime
from core.simulation import Simulation
from models.xy.model import XYModel
from utils.data_io import save_runs
from utils.temperature_ladder import make_temperature_ladder

sizes = [48, 64, 80, 96]
thermalization_sweeps = 500000
measurement_sweeps = 1000000
sample_interval = 50
exchange_interval = 10
data_path = Path("data/xy_run.npz")


def ladder():
    return make_temperature_ladder(T_min=0.78, T_max=1.02, n_T=100, Tc=0.893, tc_window=0.08,
                                   tc_fraction=0.60)


def main(batched=False):
    temperatures = ladder()
    runs = {}
    if batched:
        from core.batched import BatchedRun, BatchedSimulation
        batch = [BatchedRun(XYModel(J=1.0, delta=1.5), L, temperatures, seed=42 + L,
                            exchange_interval=exchange_interval, pt_seed=43) for L in sizes]
        runs = BatchedSimulation(batch, thermalization_sweeps, measurement_sweeps, sample_interval).run()
    else:
        for L in sizes:
            print(f"Running XY model: L = {L}")
            runtime = Runtime(model=XYModel(J=1.0, delta=1.5), L=L, temperatures=temperatures,
                              threads=16, seed=42 + L)
            pt = ParallelTempering(runtime, exchange_interval=exchange_interval, seed=43)
            runs[L] = Simulation(runtime=pt, thermalization_sweeps=thermalization_sweeps,
                                 measurement_sweeps=measurement_sweeps,
                                 sample_interval=sample_interval).run()
    save_runs(data_path, runs)
    print(f"Saved simulation to {data_path}")
    return data_path


if __name__ == "__main__":
    main()
