"""GPU parity of the persistent batched path.

* Ising: trajectories, RNG streams, betas, maps, swap counts, round trips and
  samples are bit-identical to the drop-in path (Runtime + ParallelTempering
  + Simulation) and to the CPU oracle, for one and several strips per replica.
* XY float32 fast update: trajectories and RNG streams bit-identical to the
  oracle's fast restatement; observables match it to float32 accuracy.
"""

import numpy as np
import pytest

import oracle
from gpu_helpers import device_rng_as_pairs

pytestmark = pytest.mark.gpu


def _plain_ising(L, temps, seed, interval, pt_seed, nsweeps, start="random"):
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    from models.ising.model import IsingModel
    rt = Runtime(IsingModel(start=start), L, temps, seed=seed)
    pt = ParallelTempering(rt, exchange_interval=interval, seed=pt_seed)
    for _ in range(nsweeps):
        pt.sweep()
    return rt, pt


@pytest.mark.parametrize("L,R,H,threads", [(8, 2, 8, 256), (16, 5, 4, 256), (32, 8, 32, 256), (32, 8, 8, 256),
                                           (64, 3, 16, 256), (64, 3, 32, 512), (128, 2, 16, 512)])
def test_batched_ising_matches_drop_in_path_bit_for_bit(gpu, L, R, H, threads):
    from core.batched import BatchedRun
    from models.ising.model import IsingModel
    temps = np.linspace(2.0, 2.6, R)
    seed, pt_seed, interval, n = 40 + L, 43, 3, 37
    br = BatchedRun(IsingModel(start="random"), L, temps, seed=seed, exchange_interval=interval,
                    pt_seed=pt_seed, strip_rows=H, threads=threads)
    br.run(n)
    rt, pt = _plain_ising(L, temps, seed, interval, pt_seed, n)
    np.testing.assert_array_equal(br.snapshot(), rt.snapshot())
    np.testing.assert_array_equal(device_rng_as_pairs(br.rng_states), device_rng_as_pairs(rt.rng_states))
    np.testing.assert_array_equal(br.betas.copy_to_host(), rt.betas.copy_to_host())
    d_b, d_p = br.diagnostics(), pt.diagnostics()
    for key in ("attempts", "accepted", "walker_to_temperature", "temperature_to_walker",
                "walker_round_trips", "walker_round_trip_total_sweeps"):
        np.testing.assert_array_equal(d_b[key], d_p[key], err_msg=key)
    assert d_b["exchange_rounds"] == d_p["exchange_rounds"]
    assert d_b["next_parity"] == d_p["next_parity"]
    np.testing.assert_array_equal(br.energies.copy_to_host(), pt.energies.copy_to_host())


def test_batched_ising_split_launches_continue_the_same_chain(gpu):
    from core.batched import BatchedRun
    from models.ising.model import IsingModel
    temps = np.linspace(2.1, 2.45, 6)
    one = BatchedRun(IsingModel(start="random"), 16, temps, seed=3, exchange_interval=2, strip_rows=4)
    one.run(30)
    two = BatchedRun(IsingModel(start="random"), 16, temps, seed=3, exchange_interval=2, strip_rows=4)
    two.run(7)
    two.run(23)
    np.testing.assert_array_equal(one.snapshot(), two.snapshot())
    np.testing.assert_array_equal(one.betas.copy_to_host(), two.betas.copy_to_host())


def test_launches_reuse_the_workspace_without_clearing_it(gpu):
    """Short launches in a row leave tagged halo and energy words behind; with
    the workspace clock the next launch waits only for tags above them, so the
    chain equals the one whose workspace is zeroed before every launch and the
    one run in a single launch; also with tags near the top of their range, and
    when a clock with no room left makes the launch zero the workspace first."""
    from core.batched import BatchedRun
    from models.xy.model import XYModel
    temps = np.linspace(0.8, 1.1, 6)
    chunks = (1, 1, 2, 3, 5, 10, 17, 1, 4)

    def make():
        model = XYModel(J=1.0, proposal="gaussian", sigma=1.0, start="random")
        return BatchedRun(model, 32, temps, seed=11, exchange_interval=3, pt_seed=5, strip_rows=8, threads=256)

    kept, cleared, high, single = make(), make(), make(), make()
    for i, n in enumerate(chunks):
        kept.run(n)
        assert kept._clock[0] > 0
        cleared._clock[0] = 0
        cleared.run(n)
        if i == 1:
            high._clock[0] = 0x3fff0000  # later tags near the top of their range
        if i == len(chunks) - 1:
            high._clock[0] = 0x3fffffff - 1  # no room for this launch's tags
        high.run(n)
    assert high._clock[0] == chunks[-1] + 1  # zeroed, then advanced
    single.run(sum(chunks))
    for other in (cleared, high, single):
        np.testing.assert_array_equal(kept.snapshot(), other.snapshot())
        np.testing.assert_array_equal(device_rng_as_pairs(kept.rng_states), device_rng_as_pairs(other.rng_states))
        np.testing.assert_array_equal(kept.betas.copy_to_host(), other.betas.copy_to_host())
        d_k, d_o = kept.diagnostics(), other.diagnostics()
        for key in ("attempts", "accepted", "walker_to_temperature", "walker_round_trips"):
            np.testing.assert_array_equal(d_k[key], d_o[key], err_msg=key)
    assert kept.diagnostics()["attempts"].sum() > 0


def test_single_call_entry_point_equals_the_launch_plan(gpu):
    """mc_batch_run (every argument in one call) continues a chain exactly as
    the launch plan that BatchedRun uses (mc_batch_plan_create, then
    mc_batch_plan_launch per launch), exchanges included."""
    import ctypes
    from core.batched import BatchedRun
    from mcgpu import _lib
    from mcgpu.device import as_pointer
    from models.xy.model import XYModel
    temps = np.linspace(0.8, 1.1, 6)

    def make():
        model = XYModel(J=1.0, proposal="gaussian", sigma=0.9, start="random")
        return BatchedRun(model, 32, temps, seed=21, exchange_interval=3, pt_seed=8, strip_rows=8, threads=256)

    planned, direct = make(), make()
    for n in (4, 9, 13):
        planned.run(n)
        d = direct
        partials, sw2t = d._no_samples
        _lib.call("mc_batch_run", d._stream, d.code, d.L, d.R, d.H, d.threads, as_pointer(d.state),
                  as_pointer(d.rng_states), as_pointer(d.betas), float(d.model.J), 0.0, as_pointer(d.sigmas),
                  as_pointer(d._accepted), n, d.sweeps, d.exchange_interval, as_pointer(d._parity),
                  as_pointer(d.energies), as_pointer(d.walker_to_temperature), as_pointer(d.temperature_to_walker),
                  as_pointer(d._pt_rng), as_pointer(d._swap_counts), as_pointer(d._trips), 0, 0, 0,
                  as_pointer(partials), as_pointer(sw2t), as_pointer(d._halo), as_pointer(d._energy_words),
                  as_pointer(d._error), ctypes.cast(d._clock, ctypes.c_void_p), d._error_host.pointer)
        d._advance(n)
        d.synchronize()
    assert getattr(direct, "_plan", None) is None  # the direct run never made a plan
    np.testing.assert_array_equal(planned.snapshot(), direct.snapshot())
    np.testing.assert_array_equal(device_rng_as_pairs(planned.rng_states), device_rng_as_pairs(direct.rng_states))
    np.testing.assert_array_equal(planned.betas.copy_to_host(), direct.betas.copy_to_host())
    np.testing.assert_array_equal(planned.energies.copy_to_host(), direct.energies.copy_to_host())
    d_p, d_d = planned.diagnostics(), direct.diagnostics()
    for key in ("attempts", "accepted", "walker_to_temperature", "temperature_to_walker", "walker_round_trips"):
        np.testing.assert_array_equal(d_p[key], d_d[key], err_msg=key)
    assert d_p["next_parity"] == d_d["next_parity"]
    assert d_p["attempts"].sum() > 0


def test_batched_ising_simulation_samples_equal_drop_in_simulation(gpu):
    from core.batched import BatchedRun, BatchedSimulation
    from core.simulation import Simulation
    from models.ising.model import IsingModel
    L, temps = 16, np.linspace(2.0, 2.6, 4)
    br = BatchedRun(IsingModel(start="random"), L, temps, seed=58, exchange_interval=1, pt_seed=43,
                    strip_rows=8)
    batched = BatchedSimulation([br], thermalization_sweeps=20, measurement_sweeps=60,
                                sample_interval=5).run()[L]
    rt, pt = _plain_ising(L, temps, 58, 1, 43, 0)
    plain = Simulation(pt, thermalization_sweeps=20, measurement_sweeps=60, sample_interval=5).run()
    np.testing.assert_array_equal(batched["metadata"]["sample_sweeps"], plain["metadata"]["sample_sweeps"])
    for name in ("energy", "magnetization"):
        np.testing.assert_array_equal(batched["samples"][name], plain["samples"][name], err_msg=name)


@pytest.mark.parametrize("L,R,H,threads", [(16, 3, 16, 256), (32, 4, 8, 256), (128, 2, 16, 256),
                                           (128, 2, 16, 512), (64, 3, 32, 512)])
def test_batched_xy_fast_trajectory_matches_oracle(gpu, L, R, H, threads):
    from core.batched import BatchedRun
    from models.xy.model import XYModel
    temps = np.geomspace(0.7, 1.1, R)
    model = XYModel(start="random")
    br = BatchedRun(model, L, temps, seed=170, exchange_interval=0, strip_rows=H, threads=threads)
    theta = model.initialize_state(L, R, 170).copy()
    c, s = oracle.xyfast_cache(theta)
    rng = oracle.rng_init(R * L * L, 170)
    betas = (1.0 / temps).astype(np.float32)
    samples, _ = br.run(24, sample_interval=8)
    oracle.xyfast_sweeps(theta, c, s, rng, betas, model.J, model.delta, 24)
    np.testing.assert_array_equal(br.snapshot(), theta)
    np.testing.assert_array_equal(device_rng_as_pairs(br.rng_states), rng)
    ref = oracle.xyfast_observables(c, s, model.J)
    for k, name in enumerate(("energy", "mx", "my", "twist_cos_x", "twist_sin_x")):
        np.testing.assert_allclose(samples[name][-1], ref[k], rtol=2e-6, atol=2e-3 * L, err_msg=name)
    # the cached-sincos energy agrees with the float64 reference measurement
    plain = model.measure(br.state)
    np.testing.assert_allclose(samples["energy"][-1], plain["energy"], rtol=1e-5, atol=1e-3 * L)


def test_batched_xy_fast_with_exchanges_matches_oracle_loop(gpu):
    from core.batched import BatchedRun
    from models.xy.model import XYModel
    L, R, interval = 32, 6, 5
    temps = np.geomspace(0.7, 1.1, R)
    model = XYModel(start="random")
    br = BatchedRun(model, L, temps, seed=11, exchange_interval=interval, pt_seed=12, strip_rows=8)
    br.run(40)
    theta = model.initialize_state(L, R, 11).copy()
    c, s = oracle.xyfast_cache(theta)
    rng = oracle.rng_init(R * L * L, 11)
    betas = (1.0 / temps).astype(np.float32)
    w2t = np.arange(R, dtype=np.int32)
    t2w = np.arange(R, dtype=np.int32)
    counts = np.zeros((3, R - 1), np.int64)
    trips = np.zeros((5, R), np.int64)
    trips[0] = trips[2] = -1
    trips[0, 0], trips[0, -1] = 0, R - 1
    trips[2, [0, -1]] = 0
    pt_rng = oracle.rng_init(R - 1, 12)
    parity = 0
    for sweep in range(1, 41):
        oracle.xyfast_sweeps(theta, c, s, rng, betas, model.J, model.delta, 1)
        if sweep % interval == 0:
            e = oracle.xyfast_observables(c, s, model.J)[0]
            oracle.pt_exchange(e, betas, w2t, t2w, pt_rng, counts, trips, parity, sweep)
            parity = 1 - parity
    np.testing.assert_array_equal(br.snapshot(), theta)
    np.testing.assert_array_equal(br.betas.copy_to_host(), betas)
    d = br.diagnostics()
    np.testing.assert_array_equal(d["attempts"], counts[0])
    np.testing.assert_array_equal(d["accepted"], counts[1])
    np.testing.assert_array_equal(d["walker_round_trips"], trips[3])


def test_batched_simulation_runs_several_sizes_concurrently(gpu):
    from core.batched import BatchedRun, BatchedSimulation
    from models.ising.model import IsingModel
    from models.xy.model import XYModel
    runs = [BatchedRun(IsingModel(), L, np.linspace(2.1, 2.45, 6), seed=42 + L, exchange_interval=10)
            for L in (16, 24, 32)]
    runs.append(BatchedRun(XYModel(), 32, np.linspace(0.8, 1.0, 6), seed=74, exchange_interval=10))
    out = BatchedSimulation(runs, thermalization_sweeps=50, measurement_sweeps=100,
                            sample_interval=10).run()
    assert len(out) == 4
    for run in out.values():
        for values in run["samples"].values():
            assert values.shape == (10, 6)
            assert np.all(np.isfinite(values))
        np.testing.assert_array_equal(run["metadata"]["sample_sweeps"], np.arange(60, 151, 10))


def test_campaign_shape_serialises_launches_that_do_not_fit_together(gpu):
    """main/ising.py's batched campaign: three sizes of 100 temperatures, whose
    grids together exceed the device; the launches run in waves and end
    without a watchdog error, each identical to its run alone."""
    from core.batched import BatchedRun, BatchedSimulation
    from main.ising import ladder, sizes
    from models.ising.model import IsingModel
    temps = ladder()
    make = lambda L: BatchedRun(IsingModel(), L, temps, seed=42 + L, exchange_interval=10, pt_seed=43)
    runs = [make(L) for L in sizes]
    sim = BatchedSimulation(runs, thermalization_sweeps=20, measurement_sweeps=40, sample_interval=10)
    from core.batched import max_workgroups
    shares = [r.R * r.strips / max_workgroups(r.code, r.L, r.R, r.H, r.exchange_interval, r.threads) for r in runs]
    waves = sim.waves()
    assert sorted(i for w in waves for i in w) == [0, 1, 2]
    assert all(sum(shares[i] for i in w) <= 1.0 for w in waves)
    assert len(waves) > 1 or sum(shares) <= 1.0
    out = sim.run()
    for L, run in zip(sizes, runs):
        alone = make(L)
        alone.run(60)
        np.testing.assert_array_equal(run.snapshot(), alone.snapshot())
        assert out[L]["samples"]["energy"].shape == (4, 100)


def test_watchdog_expiry_raises_and_every_kernel_exits(gpu):
    """Forced from the host: every strip's halo receives wait for tags that no
    workgroup publishes, with a short bounded wait; the launch ends, the error
    reaches BatchedRun.synchronize, and the next launch runs normally."""
    from core.batched import BatchedRun
    from mcgpu import _lib
    from models.xy.model import XYModel
    run = BatchedRun(XYModel(start="random"), 32, np.linspace(0.8, 1.0, 4), seed=3, exchange_interval=5,
                     strip_rows=8)
    _lib.call("mc_batch_test_watchdog", 1, 1 << 14)
    try:
        with pytest.raises(RuntimeError, match="watchdog"):
            run.run(20)
    finally:
        _lib.call("mc_batch_test_watchdog", 0, 0)
    run.run(20)  # the error word is cleared by the next launch
    assert np.all(np.isfinite(run.measure()["energy"]))
