"""The drop-in path's native chains against the per-launch path.

ParallelTempering.advance(n) and Runtime.advance(n) queue n sweeps (with
energies, exchanges and walker-trail rows) in one native call; Simulation
samples into a device buffer.  They must leave exactly what n calls of
sweep() through the model's per-colour update_color leave: states, RNG
streams, betas, maps, diagnostics, trails and samples, for every built-in
model.  A model whose update_color is replaced on the instance is swept
through it, as the plugin contract says.
"""

from unittest.mock import patch

import numpy as np
import pytest

from gpu_helpers import device_rng_as_pairs

pytestmark = pytest.mark.gpu


def _models():
    from models.he3.model import He3Model
    from models.ising.model import IsingModel
    from models.kawamura.model import KawamuraModel
    from models.xy.model import XYModel
    return [
        ("ising", lambda: IsingModel(start="random"), np.linspace(2.0, 2.6, 5)),
        ("xy", lambda: XYModel(start="random"), np.geomspace(0.7, 1.1, 5)),
        ("xy64", lambda: XYModel(start="random", dtype=np.float64), np.geomspace(0.7, 1.1, 5)),
        ("xy_gauss", lambda: XYModel(start="random", proposal="gaussian", sigma=np.linspace(0.5, 2, 5)),
         np.geomspace(0.7, 1.1, 5)),
        ("kawamura", lambda: KawamuraModel(start="random"), np.linspace(0.65, 0.8, 5)),
        ("he3", lambda: He3Model(pressure=20, g_d=0.001), np.linspace(0.3, 0.6, 5)),
    ]


def _pair(make, temps, L=8, seed=21, **pt_kw):
    from core.parallel_tempering import ParallelTempering
    from core.runtime import Runtime
    out = []
    for _ in range(2):
        rt = Runtime(make(), L, temps, seed=seed)
        out.append((rt, ParallelTempering(rt, exchange_interval=3, seed=seed + 1, **pt_kw)))
    return out


def _same(a, b):
    (rt_a, pt_a), (rt_b, pt_b) = a, b
    np.testing.assert_array_equal(rt_a.snapshot(), rt_b.snapshot())
    np.testing.assert_array_equal(device_rng_as_pairs(rt_a.rng_states), device_rng_as_pairs(rt_b.rng_states))
    np.testing.assert_array_equal(rt_a.betas.copy_to_host(), rt_b.betas.copy_to_host())
    np.testing.assert_array_equal(pt_a.energies.copy_to_host(), pt_b.energies.copy_to_host())
    da, db = pt_a.diagnostics(), pt_b.diagnostics()
    assert da.keys() == db.keys()
    for k in da:
        if k == "walker_history":
            for kk in da[k]:
                np.testing.assert_array_equal(da[k][kk], db[k][kk])
        else:
            np.testing.assert_array_equal(da[k], db[k])


@pytest.mark.parametrize("name,make,temps", _models(), ids=[m[0] for m in _models()])
def test_advance_equals_per_sweep_calls(gpu, name, make, temps):
    native, plain = _pair(make, temps, tracked_walkers=[0, 4, 2], history_stride=2)
    native[1].advance(4)   # ends between exchanges
    native[1].advance(11)  # ends on an exchange sweep (15)
    for _ in range(15):
        plain[1].sweep()
    assert native[1].sweeps == plain[1].sweeps == 15
    assert native[1].exchange_rounds == plain[1].exchange_rounds == 5
    _same(native, plain)
    # the energy cache after the last exchange is the PT energies in both
    for (rt, pt) in (native, plain):
        assert rt._energy.current() is pt.energies
    m_n, m_p = native[1].measure(), plain[1].measure()
    for k in m_n:
        np.testing.assert_array_equal(m_n[k], m_p[k])


def test_he3_gaussian_runtime_advance_and_counts(gpu):
    from core.parallel_tempering import ParallelTempering
    from models.he3.model import He3Model
    from models.he3.sampling import He3GaussianRuntime
    runs = []
    for _ in range(2):
        rt = He3GaussianRuntime(He3Model(pressure=20, g_d=0.001), 8, np.linspace(0.3, 0.6, 4), seed=9)
        runs.append((rt, ParallelTempering(rt, exchange_interval=2, seed=10)))
    runs[0][1].advance(9)
    for _ in range(9):
        runs[1][1].sweep()
    _same(runs[0], runs[1])
    np.testing.assert_array_equal(runs[0][0].gaussian.counts.copy_to_host(), runs[1][0].gaussian.counts.copy_to_host())
    assert runs[0][0].gaussian.sweeps == runs[1][0].gaussian.sweeps == 9


def test_simulation_device_samples_equal_host_samples(gpu):
    from core.simulation import Simulation
    from models.kawamura.model import KawamuraModel
    native, plain = _pair(lambda: KawamuraModel(start="random"), np.linspace(0.65, 0.8, 5))
    run_n = Simulation(native[1], thermalization_sweeps=7, measurement_sweeps=23, sample_interval=4).run()
    # replacing sweep on the instance forces the reference's call-per-sweep loop and host samples
    with patch.object(plain[1], "sweep", wraps=plain[1].sweep) as sweep:
        run_p = Simulation(plain[1], thermalization_sweeps=7, measurement_sweeps=23, sample_interval=4).run()
    assert sweep.call_count == 30
    np.testing.assert_array_equal(run_n["metadata"]["sample_sweeps"], run_p["metadata"]["sample_sweeps"])
    assert run_n["samples"].keys() == run_p["samples"].keys()
    for k in run_n["samples"]:
        np.testing.assert_array_equal(run_n["samples"][k], run_p["samples"][k])
    _same(native, plain)


def test_replaced_update_color_is_honoured(gpu):
    from core.runtime import Runtime
    from models.xy.model import XYModel
    rt = Runtime(XYModel(start="random"), 8, [0.8, 1.0], seed=3)
    calls = []
    original = rt.model.update_color
    with patch.object(rt.model, "update_color", side_effect=lambda *a: (calls.append(a[3]), original(*a))):
        rt.advance(2)
    assert calls == [0, 1, 2, 3, 0, 1, 2, 3]
    ref = Runtime(XYModel(start="random"), 8, [0.8, 1.0], seed=3)
    ref.advance(2)
    np.testing.assert_array_equal(rt.snapshot(), ref.snapshot())


def test_hamiltonian_change_rebinds_the_chain(gpu):
    from core.runtime import Runtime
    from models.kawamura.model import KawamuraModel
    a = Runtime(KawamuraModel(start="random"), 8, [0.7, 0.9], seed=5)
    b = Runtime(KawamuraModel(start="random"), 8, [0.7, 0.9], seed=5)
    a.sweep()
    b.sweep()
    a.model.J = np.float32(1.7)
    a.invalidate_energy_cache()
    a.sweep()
    b.model.J = np.float32(1.7)
    for color in range(4):  # the per-colour path with the new coupling
        b.model.update_color(b.state, b.rng_states, b.betas, color, b.blocks_per_grid, b.threads_per_block)
    np.testing.assert_array_equal(a.snapshot(), b.snapshot())


@pytest.mark.parametrize("L,R,dtype", [(192, 3, np.float32), (192, 2, np.float64), (12, 3, np.float32),
                                       (6, 2, np.float64), (4, 2, np.float32)])
def test_kawamura_row_sweeps_equal_four_colour_launches(gpu, L, R, dtype):
    """The row-staged sweep kernels (two launches per sweep) leave what four
    per-colour launches of update_color leave, bit for bit."""
    from core.runtime import Runtime
    from models.kawamura.model import KawamuraModel
    temps = np.linspace(0.65, 0.8, R)
    a = Runtime(KawamuraModel(start="random", dtype=dtype), L, temps, seed=31)
    b = Runtime(KawamuraModel(start="random", dtype=dtype), L, temps, seed=31)
    a.advance(5)
    for _ in range(5):
        for color in range(4):
            b.model.update_color(b.state, b.rng_states, b.betas, color, b.blocks_per_grid, b.threads_per_block)
    np.testing.assert_array_equal(a.snapshot(), b.snapshot())
    np.testing.assert_array_equal(device_rng_as_pairs(a.rng_states), device_rng_as_pairs(b.rng_states))


@pytest.mark.parametrize("L,R,dtype", [(32, 3, np.complex64), (64, 2, np.complex64), (80, 2, np.complex64),
                                       (48, 2, np.complex64), (256, 1, np.complex64), (48, 2, np.complex128),
                                       (12, 3, np.complex128), (6, 2, np.complex64), (4, 2, np.complex64),
                                       (4, 3, np.complex128)])
def test_he3_row_sweeps_equal_four_colour_launches(gpu, L, R, dtype):
    """The He3 row sweep (two launches per sweep, a workgroup owning whole
    rows; float32: the second pass's stencil from the first pass's registers
    and a neighbour through LDS) leaves what four per-colour launches of
    update_color leave, bit for bit, including the g_d dipole term, down to
    L = 4, where a row's two threads are each other's neighbours."""
    from core.runtime import Runtime
    from models.he3.model import He3Model
    temps = np.linspace(0.2, 0.4, R)

    def model():
        return He3Model(K=1.0, delta=0.15, pressure=20, reduced_temperature=1.0, g_d=0.3, dtype=dtype)

    a = Runtime(model(), L, temps, seed=31)
    b = Runtime(model(), L, temps, seed=31)
    a.advance(6)
    for _ in range(6):
        for color in range(4):
            b.model.update_color(b.state, b.rng_states, b.betas, color, b.blocks_per_grid, b.threads_per_block)
    np.testing.assert_array_equal(a.snapshot(), b.snapshot())
    np.testing.assert_array_equal(device_rng_as_pairs(a.rng_states), device_rng_as_pairs(b.rng_states))
    assert not np.array_equal(a.snapshot(), model().initialize_state(L, R, 31))  # the sweeps moved sites
