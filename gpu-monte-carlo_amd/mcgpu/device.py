"""Device arrays backed by libmcgpu's HIP allocations.

:class:`DeviceArray` offers the part of Numba's ``DeviceNDArray`` the engine
and its callers use (core/runtime.py:109-119, core/parallel_tempering.py:151,
171-188, 209-225): ``shape``, ``dtype``, ``size``, ``copy_to_host``,
``copy_to_device`` and leading-axis slicing that shares memory.
"""

import ctypes
import operator
import sys

import numpy as np

from . import _lib


class _Allocation:
    """Owns one device allocation; freed when the last view is collected."""

    __slots__ = ("ptr", "nbytes", "__weakref__")

    def __init__(self, nbytes):
        self.ptr = 0
        self.nbytes = 0
        _lib.ensure_device()
        p = ctypes.c_void_p()
        _lib.check(_lib.load().mc_malloc(ctypes.byref(p), int(nbytes)))
        self.ptr = int(p.value or 0)
        self.nbytes = int(nbytes)

    def __del__(self, _finalizing=sys.is_finalizing):
        ptr, self.ptr = self.ptr, 0
        # at interpreter exit the runtime is already shut down and module
        # globals may be None: the default argument survives that clearing
        if ptr and not _finalizing() and not _lib.closing():
            _lib.load().mc_free(ctypes.c_void_p(ptr))


class DeviceArray:
    """A C-contiguous array in GPU memory."""

    __slots__ = ("shape", "dtype", "_alloc", "_offset", "__weakref__")

    def __init__(self, shape, dtype, _alloc=None, _offset=0):
        if isinstance(shape, (int, np.integer)):
            shape = (int(shape),)
        self.shape = tuple(int(s) for s in shape)
        if any(s < 0 for s in self.shape):
            raise ValueError(f"negative dimension in shape {self.shape}")
        self.dtype = np.dtype(dtype)
        if _alloc is None:
            _alloc = _Allocation(max(self.nbytes, 1))
        self._alloc = _alloc
        self._offset = int(_offset)

    # ------------------------------------------------------------ properties
    @property
    def size(self):
        return int(np.prod(self.shape, dtype=np.int64))

    @property
    def nbytes(self):
        return self.size * self.dtype.itemsize

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def ptr(self):
        """Device address as an int."""
        return self._alloc.ptr + self._offset

    @property
    def device_ctypes_pointer(self):
        return ctypes.c_void_p(self.ptr)

    def __len__(self):
        if not self.shape:
            raise TypeError("len() of a 0-d device array")
        return self.shape[0]

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"

    # ------------------------------------------------------------- copies
    def copy_to_host(self, ary=None, stream=0):
        """Synchronous device-to-host copy; returns a new array unless `ary` is given."""
        if ary is None:
            ary = np.empty(self.shape, dtype=self.dtype)
        else:
            _check_compatible(self, ary)
            if not ary.flags.c_contiguous or not ary.flags.writeable:
                raise ValueError("destination must be a writable C-contiguous array")
        if self.nbytes:
            _lib.check(_lib.load().mc_memcpy_dtoh(ary.ctypes.data, ctypes.c_void_p(self.ptr), self.nbytes))
        return ary

    def copy_to_device(self, ary, stream=0):
        """Synchronous copy of a host array (or device array) into this array."""
        if isinstance(ary, DeviceArray):
            _check_compatible(self, ary)
            if self.nbytes:
                _lib.check(_lib.load().mc_memcpy_dtod(ctypes.c_void_p(self.ptr),
                                                      ctypes.c_void_p(ary.ptr), self.nbytes))
            return
        src = np.asarray(ary)
        _check_compatible(self, src)
        src = np.ascontiguousarray(src)
        if self.nbytes:
            _lib.check(_lib.load().mc_memcpy_htod(ctypes.c_void_p(self.ptr), src.ctypes.data, self.nbytes))

    def fill_zero(self):
        if self.nbytes:
            _lib.check(_lib.load().mc_memset(ctypes.c_void_p(self.ptr), 0, self.nbytes))

    # ------------------------------------------------------------ slicing
    def __getitem__(self, key):
        if not self.shape:
            raise IndexError("cannot index a 0-d device array")
        inner = int(np.prod(self.shape[1:], dtype=np.int64)) * self.dtype.itemsize
        if isinstance(key, slice):
            start, stop, step = key.indices(self.shape[0])
            if step != 1:
                raise NotImplementedError("device arrays support contiguous slices only")
            n = max(0, stop - start)
            return DeviceArray((n,) + self.shape[1:], self.dtype, self._alloc,
                               self._offset + start * inner)
        try:
            index = operator.index(key)
        except TypeError:
            raise NotImplementedError(
                "device arrays support integer and contiguous slice indexing on the first axis") from None
        if index < 0:
            index += self.shape[0]
        if not 0 <= index < self.shape[0]:
            raise IndexError(f"index {key} out of range for axis of length {self.shape[0]}")
        return DeviceArray(self.shape[1:], self.dtype, self._alloc, self._offset + index * inner)


def _check_compatible(dev, other):
    if np.dtype(other.dtype) != dev.dtype:
        raise TypeError(f"incompatible dtype: device array is {dev.dtype}, got {np.dtype(other.dtype)}")
    if tuple(other.shape) != dev.shape:
        if int(np.prod(other.shape, dtype=np.int64)) != dev.size or _squeeze(other.shape) != _squeeze(dev.shape):
            raise ValueError(f"incompatible shape: device array is {dev.shape}, got {tuple(other.shape)}")


def _squeeze(shape):
    return tuple(s for s in shape if s != 1)


def to_device(ary, stream=0):
    """Copy a host array to a new device array."""
    src = np.ascontiguousarray(np.asarray(ary))
    dev = DeviceArray(src.shape, src.dtype)
    dev.copy_to_device(src)
    return dev


def device_array(shape, dtype=np.float64, stream=0):
    """Allocate an uninitialised device array."""
    return DeviceArray(shape, dtype)


def device_zeros(shape, dtype=np.float64):
    dev = DeviceArray(shape, dtype)
    dev.fill_zero()
    return dev


def as_pointer(ary):
    """ctypes pointer of a DeviceArray, or NULL for None."""
    if ary is None:
        return None
    if not isinstance(ary, DeviceArray):
        raise TypeError(f"expected a DeviceArray, got {type(ary).__name__}")
    return ctypes.c_void_p(ary.ptr)


class PinnedArray:
    """A small zeroed host array in pinned memory (mc_host_alloc), the target of
    asynchronous device-to-host copies; ``host`` is a NumPy view of it."""

    def __init__(self, n, dtype):
        self.ptr = 0
        dtype = np.dtype(dtype)
        _lib.ensure_device()
        p = ctypes.c_void_p()
        nbytes = max(int(n) * dtype.itemsize, 1)
        _lib.check(_lib.load().mc_host_alloc(ctypes.byref(p), nbytes))
        self.ptr = int(p.value or 0)
        buf = (ctypes.c_char * nbytes).from_address(self.ptr)
        self.host = np.frombuffer(buf, dtype=dtype, count=int(n))

    @property
    def pointer(self):
        return ctypes.c_void_p(self.ptr)

    def __del__(self, _finalizing=sys.is_finalizing):
        ptr, self.ptr = getattr(self, "ptr", 0), 0
        self.host = None
        if ptr and not _finalizing() and not _lib.closing():
            _lib.load().mc_host_free(ctypes.c_void_p(ptr))


def synchronize():
    _lib.ensure_device()
    _lib.check(_lib.load().mc_synchronize())


def is_available():
    """True when libmcgpu loads and a HIP device is visible."""
    try:
        return _lib.device_count() > 0
    except (ImportError, OSError):
        return False
