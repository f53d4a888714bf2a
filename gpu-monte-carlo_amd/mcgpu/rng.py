"""numba-compatible xoroshiro128+ stream arrays on the GPU.

Replaces ``numba.cuda.random.create_xoroshiro128p_states`` (called at
core/runtime.py:113-115 and core/parallel_tempering.py:177-179): stream k of n
is ``jump**(k + subsequence_start)`` applied to the splitmix64 state of the
seed, records are ``{uint64 s0, uint64 s1}``.  The states are computed on the
GPU from precomputed jump matrices instead of a sequential host loop.
"""

import numpy as np

from . import _lib
from .device import DeviceArray

xoroshiro128p_dtype = np.dtype([("s0", np.uint64), ("s1", np.uint64)], align=True)

_U64 = (1 << 64) - 1


def init_xoroshiro128p_states(states, seed, subsequence_start=0, stream=0):
    if not isinstance(states, DeviceArray) or states.dtype.itemsize != 16:
        raise TypeError("states must be a DeviceArray of 16-byte xoroshiro128+ records")
    _lib.call("mc_rng_init", None, states.device_ctypes_pointer, int(states.size),
              int(seed) & _U64, int(subsequence_start))


def create_xoroshiro128p_states(n, seed, subsequence_start=0, stream=0):
    states = DeviceArray((int(n),), xoroshiro128p_dtype)
    init_xoroshiro128p_states(states, seed, subsequence_start, stream)
    return states
