"""Device arrays, streams and events through libh2hip.so's C ABI."""
import ctypes as C
import sys

import numpy as np

from ._lib import check, hip


class DeviceArray:
    """A flat device allocation holding `n` elements of `dtype`."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        check(hip().h2hip_malloc(C.byref(p), self.nbytes), "h2hip_malloc")
        self._p = p

    @classmethod
    def from_host(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    @property
    def ptr(self):
        return self._p.value or 0

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        check(hip().h2hip_memcpy(self._p, a.ctypes.data_as(C.c_void_p),
                                 self.nbytes, 0), "h2hip_memcpy")

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(hip().h2hip_memcpy(out.ctypes.data_as(C.c_void_p), self._p,
                                 self.nbytes, 1), "h2hip_memcpy")
        return out

    def fill_nan(self):
        # 0xff bytes are NaN for float64 / complex128
        check(hip().h2hip_memset(self._p, 0xff, self.nbytes), "h2hip_memset")

    def zero(self, stream=None):
        """Zeroes the array: ordered on `stream` if given, otherwise
        complete on return (so later work on any stream sees it)."""
        if stream is not None:
            check(hip().h2hip_memset_async(self._p, 0, self.nbytes,
                                           C.c_void_p(stream.handle)),
                  "h2hip_memset_async")
        else:
            check(hip().h2hip_memset(self._p, 0, self.nbytes),
                  "h2hip_memset")

    def free(self):
        if getattr(self, "_p", None) and self._p.value:
            hip().h2hip_free(self._p)
            self._p = C.c_void_p()

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        self.free()


class Stream:
    def __init__(self):
        s = C.c_void_p()
        check(hip().h2hip_stream_create(C.byref(s)), "h2hip_stream_create")
        self._s = s

    @property
    def handle(self):
        return self._s.value or 0

    def sync(self):
        check(hip().h2hip_stream_sync(self._s), "h2hip_stream_sync")

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        if getattr(self, "_s", None) and self._s.value:
            hip().h2hip_stream_destroy(self._s)
            self._s = C.c_void_p()


class Event:
    def __init__(self):
        e = C.c_void_p()
        check(hip().h2hip_event_create(C.byref(e)), "h2hip_event_create")
        self._e = e

    def record(self, stream=None):
        s = C.c_void_p(stream.handle if stream is not None else 0)
        check(hip().h2hip_event_record(self._e, s), "h2hip_event_record")

    def elapsed_ms(self, later):
        ms = C.c_float()
        check(hip().h2hip_event_elapsed(self._e, later._e, C.byref(ms)),
              "h2hip_event_elapsed")
        return ms.value

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return  # interpreter exit: the module globals are gone
        if getattr(self, "_e", None) and self._e.value:
            hip().h2hip_event_destroy(self._e)
            self._e = C.c_void_p()
