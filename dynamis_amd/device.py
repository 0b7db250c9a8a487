"""Device memory, streams and events over the C ABI (the counterparts of the reference's
GpuBuffer / Readback / GpuTimer, crates/dynamis-gpu/src/buffer.rs:6, src/readback.rs:162,
src/timing.rs:22)."""

from __future__ import annotations

import ctypes

import numpy as np

from ._native import call, lib


class DeviceBuffer:
    """An owned HIP allocation of `nbytes` bytes."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        ptr = ctypes.c_void_p()
        call("dyn_malloc", ctypes.byref(ptr), max(self.nbytes, 1))
        self.ptr = ptr.value

    @classmethod
    def from_array(cls, array: np.ndarray) -> "DeviceBuffer":
        array = np.ascontiguousarray(array)
        buf = cls(array.nbytes)
        buf.upload(array)
        return buf

    @classmethod
    def zeros(cls, nbytes: int) -> "DeviceBuffer":
        buf = cls(nbytes)
        buf.fill(0)
        return buf

    def upload(self, array: np.ndarray, offset: int = 0) -> None:
        array = np.ascontiguousarray(array)
        if offset + array.nbytes > self.nbytes:
            raise ValueError("upload beyond the end of the buffer")
        call("dyn_memcpy_htod", ctypes.c_void_p(self.ptr + offset),
             array.ctypes.data_as(ctypes.c_void_p), array.nbytes)

    def download(self, dtype, count: int | None = None, offset: int = 0) -> np.ndarray:
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        if out.nbytes + offset > self.nbytes:
            raise ValueError("download beyond the end of the buffer")
        call("dyn_memcpy_dtoh", out.ctypes.data_as(ctypes.c_void_p),
             ctypes.c_void_p(self.ptr + offset), out.nbytes)
        return out

    def fill(self, byte: int) -> None:
        call("dyn_memset", ctypes.c_void_p(self.ptr), int(byte), self.nbytes)

    def free(self) -> None:
        if self.ptr:
            call("dyn_free", ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().dyn_free(ctypes.c_void_p(self.ptr))
                self.ptr = None
        except Exception:
            pass

    @property
    def handle(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.ptr)


class Stream:
    """A non-blocking HIP stream."""

    def __init__(self):
        s = ctypes.c_void_p()
        call("dyn_stream_create", ctypes.byref(s))
        self.ptr = s.value

    def synchronize(self) -> None:
        call("dyn_stream_synchronize", ctypes.c_void_p(self.ptr))

    @property
    def handle(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(self.ptr)

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().dyn_stream_destroy(ctypes.c_void_p(self.ptr))
                self.ptr = None
        except Exception:
            pass


class Event:
    def __init__(self):
        e = ctypes.c_void_p()
        call("dyn_event_create", ctypes.byref(e))
        self.ptr = e.value

    def record(self, stream: Stream | None) -> None:
        call("dyn_event_record", ctypes.c_void_p(self.ptr),
             ctypes.c_void_p(stream.ptr if stream else None))

    def elapsed_ms(self, later: "Event") -> float:
        ms = ctypes.c_float()
        call("dyn_event_elapsed_ms", ctypes.byref(ms), ctypes.c_void_p(self.ptr),
             ctypes.c_void_p(later.ptr))
        return float(ms.value)

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().dyn_event_destroy(ctypes.c_void_p(self.ptr))
                self.ptr = None
        except Exception:
            pass


def profile_enable(on: bool) -> None:
    """Brackets every library kernel launch with events on its own stream while on."""
    call("dyn_profile_enable", 1 if on else 0)


def profile_watch(kernel: str | None) -> None:
    """Restricts the bracketing to the launches of one kernel (None: every kernel)."""
    call("dyn_profile_watch", (kernel or "").encode())


def profile_collect() -> dict[str, tuple[int, float]]:
    """Kernel name -> (launches, total ms) of the launches timed since the last collect."""
    n = ctypes.c_int()
    call("dyn_profile_collect", ctypes.byref(n))
    out = {}
    for i in range(n.value):
        name = ctypes.create_string_buffer(256)
        launches = ctypes.c_uint32()
        total = ctypes.c_double()
        call("dyn_profile_entry", i, name, 256, ctypes.byref(launches), ctypes.byref(total))
        out[name.value.decode()] = (int(launches.value), float(total.value))
    return out
