"""Helpers shared by the model plugins: argument checks and reusable
measurement output buffers (one (C, R) float64 device array per shape)."""

import numpy as np

from mcgpu import DeviceArray, as_pointer, device_array, dtype_code

_outputs = {}

STREAM = None  # default HIP stream: launches and copies are ordered


def output_buffer(channels, replicas):
    key = (int(channels), int(replicas))
    buf = _outputs.get(key)
    if buf is None:
        buf = device_array(key, dtype=np.float64)
        _outputs[key] = buf
    return buf


def require_device(array, name):
    if not isinstance(array, DeviceArray):
        raise TypeError(f"{name} must be a device array (mcgpu.DeviceArray), got {type(array).__name__}")
    return array


def indices_pointer(temperature_indices, replicas):
    if temperature_indices is None:
        return None
    require_device(temperature_indices, "temperature_indices")
    if temperature_indices.dtype != np.int32 or temperature_indices.size != replicas:
        raise ValueError("temperature_indices must be an int32 device array of length R")
    return as_pointer(temperature_indices)


def energies_pointer(energies, replicas):
    if energies is None:
        return None
    require_device(energies, "cached energies")
    if energies.dtype != np.float64 or energies.size != replicas:
        raise ValueError("energies must be a float64 device array of length R")
    return as_pointer(energies)


def real_dtype(dtype):
    """Precision of the betas a state of `dtype` is updated with (int8 Ising uses float32)."""
    dtype = np.dtype(dtype)
    if dtype in (np.float32, np.complex64, np.int8):
        return np.dtype(np.float32)
    return np.dtype(np.float64)


def check_update_args(state, rng_states, betas, color, per_site=1):
    require_device(state, "state")
    require_device(rng_states, "rng_states")
    require_device(betas, "betas")
    R, L = state.shape[0], state.shape[1]
    if state.ndim < 3 or state.shape[2] != L:
        raise ValueError(f"state must have shape (R, L, L, ...), got {state.shape}")
    if rng_states.size != R * L * L or rng_states.dtype.itemsize != 16:
        raise ValueError("rng_states must hold R*L*L xoroshiro128+ records")
    if betas.size != R:
        raise ValueError("betas must have one entry per replica")
    if betas.dtype != real_dtype(state.dtype):
        raise TypeError(f"betas dtype {betas.dtype} does not match state precision {state.dtype}")
    if not 0 <= int(color) < 4:
        raise ValueError("color must be in 0..3")
    code = dtype_code(real_dtype(state.dtype))
    return R, L, code
