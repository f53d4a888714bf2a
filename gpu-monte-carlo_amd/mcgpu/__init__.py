"""mcgpu: host side of the gfx950 Metropolis + parallel-tempering library.

``libmcgpu.so`` (built from ``csrc/*.hip``) holds every kernel; this package
loads it with ctypes and provides device arrays and RNG stream arrays.  The
drop-in modules ``core/`` and ``models/`` are written on top of it.
"""

from . import _lib
from .device import (DeviceArray, PinnedArray, as_pointer, device_array, device_zeros,
                     is_available, synchronize, to_device)
from .rng import (create_xoroshiro128p_states, init_xoroshiro128p_states,
                  xoroshiro128p_dtype)

MC_F32 = _lib.MC_F32
MC_F64 = _lib.MC_F64

__all__ = [
    "DeviceArray", "PinnedArray", "as_pointer", "device_array", "device_zeros", "is_available",
    "synchronize", "to_device", "create_xoroshiro128p_states",
    "init_xoroshiro128p_states", "xoroshiro128p_dtype", "MC_F32", "MC_F64",
    "dtype_code",
]


def dtype_code(dtype):
    """MC_F32 for float32/complex64 state, MC_F64 for float64/complex128."""
    import numpy as np
    dtype = np.dtype(dtype)
    if dtype in (np.float32, np.complex64):
        return MC_F32
    if dtype in (np.float64, np.complex128):
        return MC_F64
    raise TypeError(f"unsupported state dtype {dtype}")
