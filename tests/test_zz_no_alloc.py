"""mc_chain_energy_noalloc: the energy reduction of a native chain with its
partials in the library's static device array instead of allocated scratch.

Through the C ABI, its energies equal mc_chain_energy's bit for bit, on two
calls in a row (the second with a smaller lattice and fewer replicas) for
every built-in model, also after the static array holds the previous call's
partials; a size beyond the array's capacity is refused before any launch.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _runtimes():
    from core.runtime import Runtime
    from models.he3.model import He3Model
    from models.ising.model import IsingModel
    from models.kawamura.model import KawamuraModel
    from models.xy.model import XYModel
    return [
        (Runtime(XYModel(start="random"), 256, np.linspace(0.8, 1.1, 64), seed=7),
         Runtime(XYModel(start="random", dtype=np.float64), 16, np.linspace(0.8, 1.1, 3), seed=8)),
        (Runtime(IsingModel(start="random"), 128, np.linspace(2.0, 2.6, 32), seed=9),
         Runtime(IsingModel(start="random"), 8, np.linspace(2.0, 2.6, 2), seed=10)),
        (Runtime(KawamuraModel(start="random"), 96, np.linspace(0.65, 0.8, 16), seed=11),
         Runtime(KawamuraModel(start="random", dtype=np.float64), 12, np.linspace(0.65, 0.8, 4), seed=12)),
        (Runtime(He3Model(pressure=20, g_d=0.001), 64, np.linspace(0.3, 0.6, 8), seed=13),
         Runtime(He3Model(pressure=20, g_d=0.001, dtype=np.complex128), 6, np.linspace(0.3, 0.6, 2), seed=14)),
    ]


def test_noalloc_energy_equals_the_allocating_entry_point(gpu):
    from mcgpu import _lib, as_pointer, device_array, synchronize
    for big, small in _runtimes():
        for rt in (big, small):
            rt.advance(3)
            chain = rt.native_chain()
            ref = device_array(rt.R, np.float64)
            got = device_array(rt.R, np.float64)
            _lib.call("mc_chain_energy", chain._handle, None, as_pointer(ref))
            _lib.call("mc_chain_energy_noalloc", chain._handle, None, as_pointer(got))
            synchronize()
            r, g = ref.copy_to_host(), got.copy_to_host()
            assert np.all(np.isfinite(r))
            np.testing.assert_array_equal(g, r)


def test_noalloc_energy_refuses_sizes_beyond_its_array(gpu):
    from mcgpu import _lib, as_pointer, device_array, device_zeros
    # L = 1024 has 64 reduction blocks per replica: 1025 replicas need 65,600 partials.
    # Nothing is launched, so tiny placeholder buffers stand in for the lattice.
    tiny = device_zeros(16, np.float64)
    params = (ctypes.c_double * 2)(1.0, 1.5)
    handle = ctypes.c_void_p()
    _lib.call("mc_chain_create", 1, 0, 1024, 1025, as_pointer(tiny), as_pointer(tiny), as_pointer(tiny),
              params, 2, None, 0, ctypes.byref(handle))
    try:
        out = device_array(1025, np.float64)
        with pytest.raises(ValueError, match="mc_chain_energy_noalloc holds at most 65536"):
            _lib.call("mc_chain_energy_noalloc", handle, None, as_pointer(out))
    finally:
        _lib.load().mc_chain_destroy(handle)
