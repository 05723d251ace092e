"""Device buffers, streams and events through the HIP runtime that
libpar2gpu.so itself uses (ctypes on libamdhip64.so.7, already loaded by the
library).  Used by bench.py and the device-pointer tests, so that a process
holds one HIP runtime.
"""
from __future__ import annotations

import ctypes

from . import lib as _p2g_lib  # noqa: F401  (loads libpar2gpu.so and its HIP runtime first)

_hip = ctypes.CDLL("libamdhip64.so.7")
_P = ctypes.c_void_p
_hip.hipMalloc.argtypes = [ctypes.POINTER(_P), ctypes.c_size_t]
_hip.hipFree.argtypes = [_P]
_hip.hipMemcpy.argtypes = [_P, _P, ctypes.c_size_t, ctypes.c_int]
_hip.hipMemset.argtypes = [_P, ctypes.c_int, ctypes.c_size_t]
_hip.hipStreamCreateWithFlags.argtypes = [ctypes.POINTER(_P), ctypes.c_uint]
_hip.hipStreamDestroy.argtypes = [_P]
_hip.hipStreamSynchronize.argtypes = [_P]
_hip.hipEventCreate.argtypes = [ctypes.POINTER(_P)]
_hip.hipEventDestroy.argtypes = [_P]
_hip.hipEventRecord.argtypes = [_P, _P]
_hip.hipEventSynchronize.argtypes = [_P]
_hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), _P, _P]
_hip.hipGetErrorString.restype = ctypes.c_char_p
_hip.hipGetErrorString.argtypes = [ctypes.c_int]
_hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
_hip.hipSetDevice.argtypes = [ctypes.c_int]

H2D, D2H, D2D = 1, 2, 3


def _check(rc: int, what: str):
    if rc != 0:
        raise RuntimeError(f"{what}: {_hip.hipGetErrorString(rc).decode()}")


def device_count() -> int:
    n = ctypes.c_int(0)
    rc = _hip.hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def set_device(d: int) -> None:
    _check(_hip.hipSetDevice(d), "hipSetDevice")


def synchronize() -> None:
    _check(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")


def copy_device(dst: "DeviceBuffer", src: "DeviceBuffer", nbytes: int) -> None:
    """Device-to-device copy of nbytes (the runtime's copy kernel)."""
    if nbytes > min(dst.nbytes, src.nbytes):
        raise ValueError("copy past the end of a buffer")
    _check(_hip.hipMemcpy(dst.ptr, src.ptr, nbytes, D2D), "hipMemcpy D2D")


def download_ptr(addr: int, nbytes: int) -> bytes:
    """nbytes of device memory at a raw device address."""
    out = ctypes.create_string_buffer(max(1, nbytes))
    if nbytes:
        _check(_hip.hipMemcpy(out, _P(addr), nbytes, D2H), "hipMemcpy D2H")
    return out.raw[:nbytes]


class DeviceBuffer:
    """`nbytes` of device memory (hipMalloc: 256-byte aligned)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = _P()
        _check(_hip.hipMalloc(ctypes.byref(self.ptr), max(1, self.nbytes)), "hipMalloc")

    @classmethod
    def from_bytes(cls, data) -> "DeviceBuffer":
        b = cls(len(data))
        b.upload(data)
        return b

    @property
    def addr(self) -> int:
        return self.ptr.value or 0

    def upload(self, data, offset: int = 0) -> None:
        mv = memoryview(data).cast("B")
        n = mv.nbytes
        if offset + n > self.nbytes:
            raise ValueError("upload past the end of the buffer")
        src = (ctypes.c_char * max(1, n)).from_buffer_copy(mv) if n else None
        if n:
            _check(_hip.hipMemcpy(_P(self.addr + offset), src, n, H2D), "hipMemcpy H2D")

    def download(self, nbytes: int | None = None, offset: int = 0) -> bytes:
        n = self.nbytes - offset if nbytes is None else nbytes
        out = ctypes.create_string_buffer(max(1, n))
        if n:
            _check(_hip.hipMemcpy(out, _P(self.addr + offset), n, D2H), "hipMemcpy D2H")
        return out.raw[:n]

    def zero(self) -> None:
        _check(_hip.hipMemset(self.ptr, 0, max(1, self.nbytes)), "hipMemset")

    def free(self) -> None:
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = _P()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream at the default priority."""

    def __init__(self):
        self.handle = _P()
        _check(_hip.hipStreamCreateWithFlags(ctypes.byref(self.handle), 1), "hipStreamCreate")

    @property
    def addr(self) -> int:
        return self.handle.value or 0

    def synchronize(self) -> None:
        _check(_hip.hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def __del__(self):
        try:
            if self.handle:
                _hip.hipStreamDestroy(self.handle)
        except Exception:
            pass


class Event:
    def __init__(self):
        self.handle = _P()
        _check(_hip.hipEventCreate(ctypes.byref(self.handle)), "hipEventCreate")

    def record(self, stream: Stream | None = None) -> None:
        _check(_hip.hipEventRecord(self.handle, stream.handle if stream else None),
               "hipEventRecord")

    def elapsed_ms(self, end: "Event") -> float:
        _check(_hip.hipEventSynchronize(end.handle), "hipEventSynchronize")
        ms = ctypes.c_float(0)
        _check(_hip.hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle),
               "hipEventElapsedTime")
        return ms.value

    def __del__(self):
        try:
            if self.handle:
                _hip.hipEventDestroy(self.handle)
        except Exception:
            pass
