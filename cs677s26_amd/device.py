"""Device memory, streams and events through the C ABI (no torch, no HIP
headers on the Python side)."""
import ctypes as C

import numpy as np

from ._native import check, lib


class DeviceArray:
    """A device allocation with a numpy dtype and shape; freed on close/GC."""

    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().vj_malloc(C.byref(p), max(self.nbytes, 1)), "vj_malloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, arr, stream=None):
        arr = np.ascontiguousarray(arr)
        d = cls(arr.shape, arr.dtype)
        d.upload(arr, stream)
        return d

    def upload(self, arr, stream=None):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        assert arr.nbytes <= self.nbytes
        check(lib().vj_memcpy_htod(self.ptr, arr.ctypes.data, arr.nbytes, _s(stream)), "vj_memcpy_htod")

    def to_host(self, stream=None):
        out = np.empty(self.shape, self.dtype)
        check(lib().vj_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes, _s(stream)), "vj_memcpy_dtoh")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            lib().vj_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _s(stream):
    return None if stream is None else (stream.handle if isinstance(stream, Stream) else stream)


class Stream:
    def __init__(self):
        h = C.c_void_p()
        check(lib().vj_stream_create(C.byref(h)), "vj_stream_create")
        self.handle = h.value

    def synchronize(self):
        check(lib().vj_stream_synchronize(self.handle), "vj_stream_synchronize")

    def close(self):
        if self.handle:
            lib().vj_stream_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Event:
    def __init__(self):
        h = C.c_void_p()
        check(lib().vj_event_create(C.byref(h)), "vj_event_create")
        self.handle = h.value

    def record(self, stream=None):
        check(lib().vj_event_record(self.handle, _s(stream)), "vj_event_record")

    def synchronize(self):
        check(lib().vj_event_synchronize(self.handle), "vj_event_synchronize")

    def elapsed_ms(self, end):
        ms = C.c_float()
        check(lib().vj_event_elapsed_ms(self.handle, end.handle, C.byref(ms)), "vj_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                lib().vj_event_destroy(self.handle)
        except Exception:
            pass


def synchronize():
    check(lib().vj_device_synchronize(), "vj_device_synchronize")
