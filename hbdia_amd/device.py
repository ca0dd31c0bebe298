"""Device memory, streams, events and graphs over the C ABI of libhbdia_hip.so."""
from __future__ import annotations

import ctypes
from ctypes import byref, c_float, c_int, c_size_t, c_void_p

import numpy as np

from ._native import check_hip, hip


def device_count() -> int:
    n = c_int(0)
    check_hip(hip.hbdia_get_device_count(byref(n)))
    return n.value


def device_name() -> str:
    buf = ctypes.create_string_buffer(256)
    check_hip(hip.hbdia_device_name(buf, 256))
    return buf.value.decode()


def synchronize() -> None:
    check_hip(hip.hbdia_device_synchronize())


def mem_info() -> tuple[int, int]:
    free, total = c_size_t(0), c_size_t(0)
    check_hip(hip.hbdia_mem_info(byref(free), byref(total)))
    return free.value, total.value


class DeviceArray:
    """A flat device allocation holding `size` elements of `dtype`."""

    def __init__(self, size: int, dtype, zero: bool = False):
        self.dtype = np.dtype(dtype)
        self.size = int(size)
        self.nbytes = self.size * self.dtype.itemsize
        p = c_void_p(None)
        check_hip(hip.hbdia_malloc(byref(p), self.nbytes))
        self._ptr = p.value
        if zero and self.nbytes:
            check_hip(hip.hbdia_memset(self._ptr, 0, self.nbytes, None))
            synchronize()

    @classmethod
    def from_host(cls, arr: np.ndarray, dtype=None) -> "DeviceArray":
        a = np.ascontiguousarray(arr, dtype=dtype or arr.dtype)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    @property
    def ptr(self):
        return self._ptr

    def upload(self, arr: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.size != self.size:
            raise ValueError(f"size mismatch: {a.size} != {self.size}")
        if stream is None:
            check_hip(hip.hbdia_memcpy(self._ptr, a.ctypes.data, self.nbytes))
        else:
            check_hip(hip.hbdia_memcpy_h2d(self._ptr, a.ctypes.data, self.nbytes, stream.handle))
            stream.synchronize()

    def download(self) -> np.ndarray:
        out = np.empty(self.size, dtype=self.dtype)
        if self.nbytes:
            check_hip(hip.hbdia_memcpy(out.ctypes.data, self._ptr, self.nbytes))
        return out

    def fill_zero(self, stream=None) -> None:
        if self.nbytes:
            check_hip(hip.hbdia_memset(self._ptr, 0, self.nbytes, stream.handle if stream else None))

    def free(self) -> None:
        if self._ptr:
            check_hip(hip.hbdia_free(self._ptr))
            self._ptr = None

    def __del__(self):
        try:
            if self._ptr:
                hip.hbdia_free(self._ptr)
                self._ptr = None
        except Exception:
            pass


class PinnedArray:
    """Page-locked host memory (hipHostMalloc) viewed as a numpy array (.array);
    copies from and to it can overlap kernels."""

    def __init__(self, size: int, dtype):
        self.dtype = np.dtype(dtype)
        self.size = int(size)
        p = c_void_p(None)
        check_hip(hip.hbdia_malloc_host(byref(p), max(self.size, 1) * self.dtype.itemsize))
        self._p = p.value
        buf = (ctypes.c_char * (max(self.size, 1) * self.dtype.itemsize)).from_address(self._p)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=self.size)

    @classmethod
    def from_host(cls, arr: np.ndarray) -> "PinnedArray":
        out = cls(arr.size, arr.dtype)
        out.array[:] = arr.ravel()
        return out

    def __del__(self):
        try:
            if getattr(self, "_p", None):
                self.array = None
                hip.hbdia_free_host(self._p)
                self._p = None
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self):
        s = c_void_p(None)
        check_hip(hip.hbdia_stream_create(byref(s)))
        self.handle = s.value

    def synchronize(self) -> None:
        check_hip(hip.hbdia_stream_synchronize(self.handle))

    def wait(self, event: "Event") -> None:
        check_hip(hip.hbdia_stream_wait_event(self.handle, event.handle))

    def destroy(self) -> None:
        if self.handle:
            check_hip(hip.hbdia_stream_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            if self.handle:
                hip.hbdia_stream_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Event:
    def __init__(self):
        e = c_void_p(None)
        check_hip(hip.hbdia_event_create(byref(e)))
        self.handle = e.value

    def record(self, stream: Stream | None = None) -> None:
        check_hip(hip.hbdia_event_record(self.handle, stream.handle if stream else None))

    def synchronize(self) -> None:
        check_hip(hip.hbdia_event_synchronize(self.handle))

    def elapsed_ms(self, end: "Event") -> float:
        ms = c_float(0.0)
        check_hip(hip.hbdia_event_elapsed_ms(byref(ms), self.handle, end.handle))
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                hip.hbdia_event_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Graph:
    """A HIP graph captured from a stream and instantiated for replay."""

    def __init__(self, handle):
        self.handle = handle

    @classmethod
    def capture(cls, stream: Stream, fn) -> "Graph":
        check_hip(hip.hbdia_graph_begin_capture(stream.handle))
        try:
            fn()
        finally:
            g = c_void_p(None)
            rc = hip.hbdia_graph_end_capture(stream.handle, byref(g))
        check_hip(rc)
        return cls(g.value)

    def launch(self, stream: Stream) -> None:
        check_hip(hip.hbdia_graph_launch(self.handle, stream.handle))

    def __del__(self):
        try:
            if self.handle:
                hip.hbdia_graph_destroy(self.handle)
                self.handle = None
        except Exception:
            pass
