"""Minimal ctypes access to the HIP runtime the library uses: events for
timing work on a given stream (the library's handles launch on the stream
set with oc_*_set_stream; the default is the null stream), streams, and
device buffers to hand to entry points that take device pointers."""
from __future__ import annotations

import ctypes as C

_hip = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        for p in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
            try:
                _hip = C.CDLL(p)
                break
            except OSError:
                continue
        if _hip is None:
            raise RuntimeError("libamdhip64.so not found")
        _hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        _hip.hipEventSynchronize.argtypes = [C.c_void_p]
        _hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        _hip.hipEventDestroy.argtypes = [C.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        _hip.hipStreamCreateWithFlags.argtypes = [C.POINTER(C.c_void_p), C.c_uint]
        _hip.hipStreamDestroy.argtypes = [C.c_void_p]
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class Event:
    def __init__(self):
        self.e = C.c_void_p()
        rc = hip().hipEventCreate(C.byref(self.e))
        if rc != 0:
            raise RuntimeError(f"hipEventCreate failed ({rc})")

    def record(self, stream=None):
        rc = hip().hipEventRecord(self.e, C.c_void_p(stream or 0))
        if rc != 0:
            raise RuntimeError(f"hipEventRecord failed ({rc})")

    def synchronize(self):
        hip().hipEventSynchronize(self.e)

    def elapsed_ms(self, end: "Event") -> float:
        t = C.c_float()
        rc = hip().hipEventElapsedTime(C.byref(t), self.e, end.e)
        if rc != 0:
            raise RuntimeError(f"hipEventElapsedTime failed ({rc})")
        return t.value

    def __del__(self):
        try:
            hip().hipEventDestroy(self.e)
        except Exception:
            pass


class Stream:
    """A non-blocking HIP stream at the default priority (for handing to
    oc_*_set_stream); .handle is the hipStream_t."""

    def __init__(self):
        self.s = C.c_void_p()
        rc = hip().hipStreamCreateWithFlags(C.byref(self.s), 1)  # hipStreamNonBlocking
        if rc != 0:
            raise RuntimeError(f"hipStreamCreateWithFlags failed ({rc})")

    @property
    def handle(self) -> int:
        return self.s.value or 0

    def synchronize(self):
        hip().hipStreamSynchronize(self.s)

    def __del__(self):
        try:
            hip().hipStreamDestroy(self.s)
        except Exception:
            pass


def device_synchronize():
    hip().hipDeviceSynchronize()


def device_copy(nbytes: int) -> int:
    """One device-to-device hipMemcpy of nbytes between two fresh buffers
    (the runtime runs it as a copy kernel that reads and writes nbytes once;
    the source is filled first by hipMemset).  Returns nbytes."""
    h = hip()
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        rc = h.hipMalloc(C.byref(p), C.c_size_t(nbytes))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")
    try:
        for call, args in (("hipMemset", (a, C.c_int(1), C.c_size_t(nbytes))),
                           ("hipDeviceSynchronize", ()),
                           ("hipMemcpy", (b, a, C.c_size_t(nbytes), C.c_int(3))),  # DeviceToDevice
                           ("hipDeviceSynchronize", ())):
            rc = getattr(h, call)(*args)
            if rc != 0:
                raise RuntimeError(f"{call} failed ({rc})")
    finally:
        h.hipFree(a)
        h.hipFree(b)
    return nbytes


class DeviceBuffer:
    """nbytes of device memory from hipMalloc (for handing a device pointer,
    such as a vertex buffer, to the library); .ptr is its address."""

    def __init__(self, nbytes: int):
        self.p = C.c_void_p()
        self.nbytes = int(nbytes)
        rc = hip().hipMalloc(C.byref(self.p), C.c_size_t(self.nbytes))
        if rc != 0:
            raise RuntimeError(f"hipMalloc failed ({rc})")

    @property
    def ptr(self) -> int:
        return self.p.value or 0

    def fill_zero(self):
        rc = hip().hipMemset(self.p, C.c_int(0), C.c_size_t(self.nbytes))
        if rc != 0:
            raise RuntimeError(f"hipMemset failed ({rc})")

    def to_host(self, out) -> None:
        """Copy the whole buffer into the contiguous numpy array out
        (synchronous hipMemcpy, device to host)."""
        assert out.nbytes == self.nbytes and out.flags["C_CONTIGUOUS"]
        rc = hip().hipMemcpy(C.c_void_p(out.ctypes.data), self.p, C.c_size_t(self.nbytes), C.c_int(2))
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed ({rc})")

    def __del__(self):
        try:
            hip().hipFree(self.p)
        except Exception:
            pass
