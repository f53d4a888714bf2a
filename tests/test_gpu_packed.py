"""GPU parity of the packed launch (several ensembles in one launch).

* Every ensemble of a pack -- Ising, XY uniform and XY Gaussian, mixed in one
  launch -- ends bit-identical to the same ensemble run alone by the resident
  BatchedRun: lattice, RNG streams, betas, walker maps, PT energies, swap
  counts, round trips and acceptance counters; Ising samples exactly, XY
  samples to float64 rounding (strips are added in another grouping).  This
  holds for small tasks (many strips, one replica each), large tasks (whole
  replicas, several per task) and 512-thread workgroups.
* Graph replays continue the chains exactly as eager launches do, with eager
  work between replays, and PackedSimulation with and without a graph agree.
* At the finite-size-scaling shapes (48 temperatures, Ising L = 16..256 and
  XY L = 16..256 in one launch) the largest Ising size equals the drop-in
  path and an XY size equals the oracle's fast-path loop.
"""

import numpy as np
import pytest

import oracle
from gpu_helpers import device_rng_as_pairs

pytestmark = pytest.mark.gpu

PT_KEYS = ("attempts", "accepted", "walker_to_temperature", "temperature_to_walker", "walker_round_trips",
           "walker_round_trip_total_sweeps")


def _members():
    from models.ising.model import IsingModel
    from models.xy.model import XYModel
    return [
        (IsingModel(start="random"), 8, np.linspace(2.0, 2.6, 4), 5),
        (IsingModel(start="random"), 16, np.linspace(2.0, 2.6, 6), 3),
        (IsingModel(start="random"), 24, np.linspace(2.1, 2.45, 4), 4),
        (XYModel(start="random"), 16, np.geomspace(0.7, 1.1, 4), 5),
        (XYModel(start="random", proposal="gaussian", sigma=1.2), 32, np.geomspace(0.7, 1.1, 8), 5),
    ]


def _ensembles(cls, **kw):
    out = []
    for i, (model, L, temps, interval) in enumerate(_members()):
        e = cls(model, L, temps, seed=100 + L + i, exchange_interval=interval, pt_seed=7 + i, **kw)
        if e.sigmas is not None:
            e.set_sigmas(np.linspace(0.6, 2.0, e.R))
        out.append(e)
    return out


def _assert_same_ensemble(a, b, sa=None, sb=None, exact_samples=True):
    np.testing.assert_array_equal(a.snapshot(), b.snapshot())
    np.testing.assert_array_equal(device_rng_as_pairs(a.rng_states), device_rng_as_pairs(b.rng_states))
    np.testing.assert_array_equal(a.betas.copy_to_host(), b.betas.copy_to_host())
    np.testing.assert_array_equal(a.energies.copy_to_host(), b.energies.copy_to_host())
    np.testing.assert_array_equal(a._accepted.copy_to_host(), b._accepted.copy_to_host())
    da, db = a.diagnostics(), b.diagnostics()
    for key in PT_KEYS:
        np.testing.assert_array_equal(da[key], db[key], err_msg=key)
    assert da["next_parity"] == db["next_parity"] and da["exchange_rounds"] == db["exchange_rounds"]
    if sa is not None:
        (xa, wa), (xb, wb) = sa, sb
        np.testing.assert_array_equal(wa, wb)
        for name in xa:
            if exact_samples:
                np.testing.assert_array_equal(xa[name], xb[name], err_msg=name)
            else:
                np.testing.assert_allclose(xa[name], xb[name], rtol=1e-12, atol=1e-9, err_msg=name)


@pytest.mark.parametrize("threads,task_sites", [(256, 64), (256, 4096), (512, 1024)])
def test_pack_matches_single_ensemble_launches(gpu, threads, task_sites):
    from core.batched import BatchedRun, Ensemble
    from core.packed import PackedRun
    packed = _ensembles(Ensemble)
    pr = PackedRun(packed, threads=threads, task_sites=task_sites)
    shapes = pr.describe()
    samples = pr.run(40, sample_interval=4)
    alone = _ensembles(BatchedRun)
    for e, b, s, d in zip(packed, alone, samples, shapes):
        sb = b.run(40, sample_interval=4)
        _assert_same_ensemble(e, b, s, sb, exact_samples=e.code == 0)
    assert any(d["replicas_per_task"] > 1 for d in shapes) or task_sites < 4096
    if task_sites == 64:
        assert sum(d["strip_rows"] < d["L"] for d in shapes) >= 4  # several strips per replica


def test_graph_replays_continue_the_eager_chain(gpu):
    from core.batched import Ensemble
    from core.packed import PackedRun
    eager = _ensembles(Ensemble)
    pe = PackedRun(eager, threads=256, task_sites=512)
    graphed = _ensembles(Ensemble)
    pg = PackedRun(graphed, threads=256, task_sites=512)
    graph = pg.capture(20, 3, sample_interval=5)
    eager_samples = []
    for k in range(3):
        eager_samples.append(pe.run(20, sample_interval=5))
        graph.replay()
        pg.synchronize()
        for e, g in zip(eager, graphed):
            np.testing.assert_array_equal(e.snapshot(), g.snapshot(), err_msg=f"replay {k}")
            np.testing.assert_array_equal(device_rng_as_pairs(e.rng_states), device_rng_as_pairs(g.rng_states))
            np.testing.assert_array_equal(e.walker_to_temperature.copy_to_host(),
                                          g.walker_to_temperature.copy_to_host())
        graphed[0].measure()  # eager work on the ensembles between replays
        pg.run(0)  # an eager launch with its own (placeholder) sample buffers: the graph keeps its own
    with pytest.raises(ValueError):
        graph.replay()
    got = graph.collect()
    for i, (x, w) in enumerate(got):
        for name in x:
            ref = np.concatenate([eager_samples[k][i][0][name] for k in range(3)])
            np.testing.assert_array_equal(x[name], ref, err_msg=name)
        np.testing.assert_array_equal(w, np.arange(5, 61, 5))
    for e, g in zip(eager, graphed):
        _assert_same_ensemble(e, g)


def test_packed_simulation_with_and_without_graph(gpu):
    from core.batched import Ensemble
    from core.packed import PackedRun, PackedSimulation
    from utils.pt_analysis import PTAnalysis
    runs = []
    for chunk in (None, 30):
        pr = PackedRun(_ensembles(Ensemble), threads=256, task_sites=1024)
        runs.append(PackedSimulation(pr, thermalization_sweeps=30, measurement_sweeps=60, sample_interval=10,
                                     chunk=chunk).run())
    a, b = runs
    assert sorted(a, key=str) == sorted(b, key=str)
    for key in a:
        np.testing.assert_array_equal(a[key]["metadata"]["sample_sweeps"], np.arange(40, 91, 10))
        for name in a[key]["samples"]:
            np.testing.assert_array_equal(a[key]["samples"][name], b[key]["samples"][name])
        assert a[key]["metadata"]["engine"] == "packed"
    summary = PTAnalysis(a).summary()
    assert len(summary) == len(a)
    for res in summary.values():
        assert all(n > 0 for n in res["attempts"])


def test_finite_size_set_in_one_launch(gpu):
    """Ising L = 16..256 and XY L = 16..256 with 48 temperatures each, one launch."""
    from core.batched import Ensemble
    from core.packed import PackedRun
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from models.ising.model import IsingModel
    from models.xy.model import XYModel
    ising_T = np.linspace(2.1, 2.45, 48)
    xy_T = np.linspace(0.8, 1.0, 48)
    ens = [Ensemble(IsingModel(start="random"), L, ising_T, seed=42 + L, exchange_interval=10, pt_seed=43 + L)
           for L in (16, 24, 32, 48, 64, 96, 128, 256)]
    ens += [Ensemble(XYModel(start="random"), L, xy_T, seed=42 + L, exchange_interval=10, pt_seed=43 + L)
            for L in (16, 32, 64, 128, 256)]
    pr = PackedRun(ens, threads=256, task_sites=2048)
    n = 30
    pr.run(n)
    # Ising L = 256 against the drop-in path
    big = ens[7]
    rt = Runtime(IsingModel(start="random"), 256, ising_T, seed=42 + 256)
    pt = ParallelTempering(rt, exchange_interval=10, seed=43 + 256)
    for _ in range(n):
        pt.sweep()
    np.testing.assert_array_equal(big.snapshot(), rt.snapshot())
    d_b, d_p = pr.diagnostics(7), pt.diagnostics()
    for key in PT_KEYS:
        np.testing.assert_array_equal(d_b[key], d_p[key], err_msg=key)
    # XY L = 64 against the oracle's fast-path loop with exchanges
    xy = ens[10]
    R, L = 48, 64
    theta = XYModel(start="random").initialize_state(L, R, 42 + L).copy()
    c, s = oracle.xyfast_cache(theta)
    rng = oracle.rng_init(R * L * L, 42 + L)
    betas = (1.0 / xy_T).astype(np.float32)
    w2t = np.arange(R, dtype=np.int32)
    t2w = np.arange(R, dtype=np.int32)
    counts = np.zeros((3, R - 1), np.int64)
    trips = np.zeros((5, R), np.int64)
    trips[0] = trips[2] = -1
    trips[0, 0], trips[0, -1] = 0, R - 1
    trips[2, [0, -1]] = 0
    pt_rng = oracle.rng_init(R - 1, 43 + L)
    parity = 0
    for sweep in range(1, n + 1):
        oracle.xyfast_sweeps(theta, c, s, rng, betas, 1.0, 1.5, 1)
        if sweep % 10 == 0:
            e = oracle.xyfast_observables(c, s, 1.0)[0]
            oracle.pt_exchange(e, betas, w2t, t2w, pt_rng, counts, trips, parity, sweep)
            parity = 1 - parity
    np.testing.assert_array_equal(xy.snapshot(), theta)
    np.testing.assert_array_equal(xy.betas.copy_to_host(), betas)
    np.testing.assert_array_equal(xy.diagnostics()["accepted"], counts[1])


def test_fss_study_small(gpu):
    """The finite-size-scaling comparison on small sizes: Ising samples of the
    packed launch equal the plain path's; XY helicity and Binder curves and
    their crossings agree within errors."""
    from main.fss_study import run
    rep = run(ising_sizes=(8, 12), xy_sizes=(8, 16), R=8, thermalization=2000, measurement=10000, interval=10)
    assert rep["ising_samples_identical"]
    for key, v in rep["xy"]["z"].items():
        if "mean_z2" in v:
            assert v["mean_z2"] < 3.0 and v["max_abs_z"] < 4.5, (key, v)
        else:
            assert abs(v["z"]) < 4.5, (key, v)
