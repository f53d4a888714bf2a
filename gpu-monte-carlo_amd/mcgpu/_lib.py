"""ctypes binding of libmcgpu.so (the gfx950 HIP library, see include/mcgpu.h).

There is no fallback: if the shared library is missing or no HIP device is
visible, the first GPU call raises.

One HIP runtime per process.  libmcgpu binds to the HIP runtime already in
the process when there is one, else to ROCm's (its RUNPATH):
* torch imported first: torch has loaded its bundled libamdhip64, whose
  soname (libamdhip64.so.7) satisfies libmcgpu's dependency, so both share
  one runtime and device pointers and streams move freely between torch
  tensors and this library;
* torch not imported: libmcgpu uses ROCm's runtime, the one rocprofv3
  instruments.  torch's libraries name their runtime by its unversioned file
  name, which the dynamic linker does not match against an already loaded
  libamdhip64.so.7, so importing torch after the first GPU call would load a
  second runtime; ``runtime_paths()`` lists the runtimes mapped, and the
  torch-facing modules refuse to run with two.

At interpreter exit an atexit hook waits for queued GPU work and then stops
every HIP call from finalisers (device arrays, events, streams collected
during shutdown), so nothing reaches a runtime that is being torn down.
"""

import atexit
import ctypes
import os
import sys
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmcgpu.so"

MC_OK, MC_ERR_VALUE, MC_ERR_HIP, MC_ERR_NOTIMPL, MC_ERR_TYPE = 0, 1, 2, 3, 4
MC_F32, MC_F64 = 0, 1

P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
U64 = ctypes.c_uint64
F32 = ctypes.c_float
F64 = ctypes.c_double
SZ = ctypes.c_size_t
INT = ctypes.c_int

# name -> argument types (every function returns int status)
SIGNATURES = {
    "mc_device_count": [ctypes.POINTER(INT)],
    "mc_init": [INT],
    "mc_malloc": [ctypes.POINTER(P), SZ],
    "mc_free": [P],
    "mc_host_alloc": [ctypes.POINTER(P), SZ],
    "mc_host_free": [P],
    "mc_memcpy_htod": [P, P, SZ],
    "mc_memcpy_dtoh": [P, P, SZ],
    "mc_memcpy_dtod": [P, P, SZ],
    "mc_memset": [P, INT, SZ],
    "mc_synchronize": [],
    "mc_stream_create": [ctypes.POINTER(P)],
    "mc_stream_destroy": [P],
    "mc_stream_synchronize": [P],
    "mc_event_create": [ctypes.POINTER(P)],
    "mc_event_destroy": [P],
    "mc_event_record": [P, P],
    "mc_event_elapsed_ms": [P, P, ctypes.POINTER(F32)],
    "mc_device_info": [ctypes.c_char_p, INT, ctypes.POINTER(INT)],
    "mc_rng_init": [P, P, I64, U64, U64],
    "mc_ising_update_color": [P, P, P, P, INT, INT, INT, F32],
    "mc_ising_measure": [P, P, INT, INT, F32, P, P, P],
    "mc_ising_energy": [P, P, INT, INT, F32, P],
    "mc_xy_update_color": [P, INT, P, P, P, INT, INT, INT, F64, F64],
    "mc_xy_gauss_update_color": [P, INT, P, P, P, INT, INT, INT, F64, P, P, INT, P],
    "mc_xy_measure": [P, INT, P, INT, INT, F64, P, P, P],
    "mc_xy_energy": [P, INT, P, INT, INT, F64, P],
    "mc_kawamura_update_color": [P, INT, P, P, P, INT, INT, INT, F64, F64],
    "mc_kawamura_sweep": [P, INT, P, P, P, INT, INT, F64, F64],
    "mc_kawamura_measure": [P, INT, P, INT, INT, F64, P, P, P],
    "mc_kawamura_energy": [P, INT, P, INT, INT, F64, P],
    "mc_he3_update_color": [P, INT, P, P, P, INT, INT, INT, P, F64, F64, F64],
    "mc_he3_sweep": [P, INT, P, P, P, INT, INT, P, F64, F64, F64],
    "mc_he3_gaussian_update_color": [P, INT, P, P, P, INT, INT, INT, P, F64, F64, F64, P, F64],
    "mc_he3_measure": [P, INT, P, INT, INT, P, F64, F64, P, P, P],
    "mc_he3_energy": [P, INT, P, INT, INT, P, F64, F64, P],
    "mc_pt_exchange": [P, INT, P, P, P, P, P, P, P, INT, INT, I64],
    "mc_pt_record_walkers": [P, P, P, INT, P],
    "mc_chain_create": [INT, INT, INT, INT, P, P, P, P, INT, P, INT, ctypes.POINTER(P)],
    "mc_chain_destroy": [P],
    "mc_chain_channels": [P, ctypes.POINTER(INT)],
    "mc_chain_attach_exchange": [P, P, P, P, P, P, P],
    "mc_chain_sweeps": [P, P, ctypes.c_longlong],
    "mc_chain_energy": [P, P, P],
    "mc_chain_energy_noalloc": [P, P, P],
    "mc_chain_measure": [P, P, P, P, P],
    "mc_chain_run": [P, P, ctypes.c_longlong, ctypes.c_longlong, INT, INT, ctypes.c_longlong, INT, P, INT, P,
                     ctypes.c_longlong, ctypes.c_longlong, ctypes.POINTER(INT), ctypes.POINTER(ctypes.c_longlong),
                     ctypes.POINTER(ctypes.c_longlong)],
    "mc_xy_twist_terms": [P, INT, P, INT, INT, P, P],
    "mc_xy_twist_grad": [P, INT, P, INT, INT, P, P, P, P, P],
    "mc_xy_twist_hvp": [P, INT, P, INT, INT, P, P, P, P, P, P],
    "mc_kawamura_twist_terms": [P, INT, P, INT, INT, P, P],
    "mc_kawamura_twist_grad": [P, INT, P, INT, INT, P, P, P, P, P],
    "mc_batch_workspace_size": [INT, INT, INT, INT, ctypes.POINTER(SZ), ctypes.POINTER(SZ)],
    "mc_batch_max_workgroups": [INT, INT, INT, INT, INT, INT, ctypes.POINTER(INT)],
    "mc_batch_test_watchdog": [INT, ctypes.c_uint],
    "mc_batch_set_cooperative": [INT],
    "mc_batch_run": [P, INT, INT, INT, INT, INT, P, P, P, F32, F32, P, P, INT, I64, INT, P, P, P, P, P, P, P,
                     INT, I64, INT, P, P, P, P, P, P, P],
    "mc_batch_plan_create": [P, INT, INT, INT, INT, INT, P, P, P, F32, F32, P, P, INT, P, P, P, P, P, P, P, P, P,
                             P, P, P, ctypes.POINTER(P)],
    "mc_batch_plan_launch": [P, INT, I64, INT, I64, INT, P, P],
    "mc_batch_plan_destroy": [P],
    "mc_pack_lds_bytes": [INT, INT, INT, INT, INT, ctypes.POINTER(SZ)],
    "mc_pack_occupancy": [INT, SZ, ctypes.POINTER(INT)],
    "mc_pack_create": [INT, ctypes.POINTER(P)],
    "mc_pack_destroy": [P],
    "mc_pack_add_group": [P, INT, INT, INT, INT, INT, INT, F32, F32, P, P, P, P, P, P, P, P, P, P, P, P,
                          ctypes.POINTER(INT)],
    "mc_pack_set_work": [P, P, INT],
    "mc_pack_set_sweeps": [P, I64],
    "mc_pack_sweeps": [P, ctypes.POINTER(I64)],
    "mc_pack_error": [P, ctypes.POINTER(INT)],
    "mc_pack_run": [P, P, INT, INT, I64, INT, P, P],
    "mc_pack_capture": [P, P, INT, INT, I64, INT, P, P, ctypes.POINTER(P)],
    "mc_graph_launch": [P, P],
    "mc_graph_destroy": [P],
}

_lock = threading.RLock()
_lib = None
_initialised = False
_closing = False


class MCError(RuntimeError):
    """HIP runtime failure inside libmcgpu."""


def runtime_paths():
    """Real paths of the HIP runtimes (libamdhip64) mapped into this process."""
    found = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                path = line.rstrip("\n").split(None, 5)[-1]
                if "/libamdhip64.so" in path:
                    found.add(os.path.realpath(path))
    except OSError:
        pass
    return sorted(found)


def load():
    """Load libmcgpu.so and declare its signatures (no device is touched)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python -m mcgpu.build` "
                    "(or __graft_entry__.build()); there is no CPU fallback")
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, args in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = INT
            lib.mc_last_error.argtypes = []
            lib.mc_last_error.restype = ctypes.c_char_p
            lib.mc_version.argtypes = []
            lib.mc_version.restype = ctypes.c_char_p
            _lib = lib
    return _lib


def closing():
    """True once the interpreter has started to shut down (no more HIP calls)."""
    return _closing or sys.is_finalizing()


@atexit.register
def _shutdown():
    global _closing
    if _initialised and not _closing:
        try:
            _lib.mc_synchronize()
        except Exception:
            pass
    _closing = True


def check(status):
    if status == MC_OK:
        return
    message = load().mc_last_error().decode(errors="replace")
    if status == MC_ERR_VALUE:
        raise ValueError(message)
    if status == MC_ERR_TYPE:
        raise TypeError(message)
    if status == MC_ERR_NOTIMPL:
        raise NotImplementedError(message)
    raise MCError(message)


def device_count():
    n = INT(0)
    check(load().mc_device_count(ctypes.byref(n)))
    return int(n.value)


def ensure_device(device=0):
    """Select the HIP device once per process; raises if none is visible."""
    global _initialised
    if _initialised:
        return
    lib = load()
    with _lock:
        if not _initialised:
            if device_count() < 1:
                raise MCError("no HIP device is visible; libmcgpu requires an AMD GPU")
            check(lib.mc_init(int(device)))
            _initialised = True


def call(name, *args):
    """Call a GPU entry point, initialising the device first."""
    ensure_device()
    check(getattr(load(), name)(*args))
