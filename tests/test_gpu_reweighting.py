"""Float64 multiple-histogram reweighting of the Ising benchmark ladder
(L = 32, 8 temperatures 2.0-2.6, random start, exchanges every sweep,
2,000 + 10,000 sweeps) run on the GPU drop-in path: the energy reweighted to
the midpoints between ladder temperatures equals Kaufman's exact energy of
the periodic 32 x 32 lattice within jackknife errors."""

import numpy as np
import pytest

from test_oracle_statistics import kaufman_energy

pytestmark = pytest.mark.gpu


def test_reweighted_ising_ladder_matches_kaufman_between_temperatures(gpu):
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from core.simulation import Simulation
    from models.ising.model import IsingModel
    from utils.reweighting import reweight_run
    L, R = 32, 8
    temps = np.linspace(2.0, 2.6, R)
    rt = Runtime(IsingModel(start="random"), L, temps, seed=42 + L)
    pt = ParallelTempering(rt, exchange_interval=1, seed=43)
    run = Simulation(pt, thermalization_sweeps=2000, measurement_sweeps=10000, sample_interval=1).run()
    mid = 0.5 * (temps[1:] + temps[:-1])
    out = reweight_run(run, 1.0 / mid, block_size=500)
    exact = np.array([kaufman_energy(L, 1.0 / t) for t in mid])
    z = (out["energy"] - exact) / out["energy_err"]
    assert np.all(np.abs(z) < 5), (out["energy"] / L**2, exact / L**2, out["energy_err"] / L**2)
    assert np.all(out["specific_heat"] > 0)
