"""HIP-event timing of work queued on one stream."""

import ctypes
import sys

from . import _lib


class EventTimer:
    """Records a start and a stop event on `stream`; elapsed_ms() waits for the stop."""

    def __init__(self, stream=None):
        _lib.ensure_device()
        self.stream = stream
        self._start, self._stop = ctypes.c_void_p(), ctypes.c_void_p()
        _lib.check(_lib.load().mc_event_create(ctypes.byref(self._start)))
        _lib.check(_lib.load().mc_event_create(ctypes.byref(self._stop)))

    def start(self):
        _lib.check(_lib.load().mc_event_record(self._start, self.stream))

    def stop(self):
        _lib.check(_lib.load().mc_event_record(self._stop, self.stream))

    def elapsed_ms(self):
        ms = ctypes.c_float()
        _lib.check(_lib.load().mc_event_elapsed_ms(self._start, self._stop, ctypes.byref(ms)))
        return float(ms.value)

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing() or _lib.closing():
            return
        _lib.load().mc_event_destroy(self._start)
        _lib.load().mc_event_destroy(self._stop)


def device_info():
    _lib.ensure_device()
    buf = ctypes.create_string_buffer(256)
    cus = ctypes.c_int()
    _lib.check(_lib.load().mc_device_info(buf, 256, ctypes.byref(cus)))
    return buf.value.decode(errors="replace"), int(cus.value)
