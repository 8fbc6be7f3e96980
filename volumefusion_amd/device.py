"""Device memory, streams and event timing over the library's own C ABI
(no GPU framework needed on the host)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class Stream:
    """A non-blocking HIP stream (or the default stream when `default`)."""

    def __init__(self, default: bool = False):
        self.handle = C.c_void_p(None)
        if not default:
            L.call("vf_stream_create", C.byref(self.handle))
        self._owned = not default

    @property
    def ptr(self):
        return self.handle

    def synchronize(self) -> None:
        L.call("vf_stream_synchronize", self.handle)

    def close(self) -> None:
        if self._owned and self.handle:
            L.load().vf_stream_destroy(self.handle)
            self.handle = C.c_void_p(None)
            self._owned = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


DEFAULT_STREAM = None


def default_stream() -> Stream:
    global DEFAULT_STREAM
    if DEFAULT_STREAM is None:
        DEFAULT_STREAM = Stream(default=True)
    return DEFAULT_STREAM


class DeviceArray:
    """A typed device allocation with numpy-style shape."""

    def __init__(self, shape, dtype, stream: Stream | None = None):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        self._ptr = C.c_void_p(None)
        self.stream = stream or default_stream()
        if self.nbytes:
            L.call("vf_malloc", C.byref(self._ptr), self.nbytes)

    @classmethod
    def from_numpy(cls, a: np.ndarray, stream: Stream | None = None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, stream)
        d.upload(a)
        return d

    @property
    def ptr(self) -> int:
        return self._ptr.value or 0

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"upload of {a.nbytes} B into {self.nbytes} B")
        if self.nbytes:
            L.call("vf_memcpy_h2d", self._ptr, a.ctypes.data_as(C.c_void_p), self.nbytes, self.stream.ptr)
            self.stream.synchronize()

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            L.call("vf_memcpy_d2h", out.ctypes.data_as(C.c_void_p), self._ptr, self.nbytes, self.stream.ptr)
        return out

    def fill(self, value) -> None:
        if not self.nbytes:
            return
        if self.dtype == np.float32:
            L.call("vf_fill_f32", self._ptr, float(value), self.nbytes // 4, self.stream.ptr)
        elif value == 0:
            L.call("vf_memset", self._ptr, 0, self.nbytes, self.stream.ptr)
        else:
            self.upload(np.full(self.shape, value, self.dtype))

    def copy_from(self, other: "DeviceArray") -> None:
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        L.call("vf_memcpy_d2d", self._ptr, other._ptr, self.nbytes, self.stream.ptr)

    def free(self) -> None:
        if self._ptr:
            L.load().vf_free(self._ptr)
            self._ptr = C.c_void_p(None)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """Page-locked host memory (vf_host_alloc) viewed as a numpy array."""

    def __init__(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        self.nbytes = max(int(a.nbytes), 1)
        self._ptr = C.c_void_p(None)
        L.call("vf_host_alloc", C.byref(self._ptr), self.nbytes)
        buf = (C.c_char * self.nbytes).from_address(self._ptr.value)
        self.array = np.frombuffer(buf, dtype=a.dtype, count=a.size).reshape(a.shape)
        self.array[...] = a

    @property
    def ptr(self) -> int:
        return int(self._ptr.value or 0)

    def free(self) -> None:
        if self._ptr:
            self.array = None
            L.load().vf_host_free(self._ptr)
            self._ptr = C.c_void_p(None)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = C.c_void_p(None)
        L.call("vf_event_create", C.byref(self.handle))

    def record(self, stream: Stream) -> None:
        L.call("vf_event_record", self.handle, stream.ptr)

    def synchronize(self) -> None:
        L.call("vf_event_synchronize", self.handle)

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        L.call("vf_event_elapsed_ms", C.byref(ms), self.handle, later.handle)
        return float(ms.value)

    def __del__(self):
        try:
            if self.handle:
                L.load().vf_event_destroy(self.handle)
        except Exception:
            pass


def device_count() -> int:
    n = C.c_int(0)
    st = L.load().vf_device_count(C.byref(n))
    return int(n.value) if st == 0 else 0


def synchronize() -> None:
    L.call("vf_device_synchronize")
