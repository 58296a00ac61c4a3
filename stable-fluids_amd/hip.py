"""Minimal ctypes binding of the HIP runtime (libamdhip64) used by the host-side
mirror, the tests and the benchmark: device buffers, copies, streams and events.
Plumbing only; all simulation work runs in libstable-fluids.so.
"""
import ctypes as C
import os

import numpy as np

_H2D, _D2H, _D2D = 1, 2, 3
_lib = None


def lib():
    global _lib
    if _lib is None:
        for name in ("libamdhip64.so", "libamdhip64.so.7", os.path.join("/opt/rocm/lib", "libamdhip64.so")):
            try:
                _lib = C.CDLL(name)
                break
            except OSError:
                continue
        if _lib is None:
            raise RuntimeError("HIP runtime (libamdhip64.so) not found")
        _lib.hipGetErrorString.restype = C.c_char_p
        _lib.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    return _lib


def check(code, what):
    if code != 0:
        msg = lib().hipGetErrorString(code).decode()
        raise RuntimeError(f"{what} failed: {code} ({msg})")


def device_count():
    n = C.c_int(0)
    if lib().hipGetDeviceCount(C.byref(n)) != 0:
        return 0
    return n.value


_ATTR_MULTIPROCESSOR_COUNT = 63  # hipDeviceAttributeMultiprocessorCount (hip_runtime_api.h)


def compute_units(device=0):
    """Compute units of a device (hipDeviceGetAttribute)."""
    n = C.c_int(0)
    check(lib().hipDeviceGetAttribute(C.byref(n), _ATTR_MULTIPROCESSOR_COUNT, device), "hipDeviceGetAttribute")
    return n.value


class DeviceBuffer:
    """A hipMalloc'd device array of float32 (or raw bytes)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        check(lib().hipMalloc(C.byref(self.ptr), C.c_size_t(max(1, self.nbytes))), "hipMalloc")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        check(lib().hipMemcpy(buf.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), _H2D), "hipMemcpy H2D")
        return buf

    def free(self):
        if self.ptr:
            lib().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def value(self):
        return self.ptr.value


def to_host(dev_ptr, count, dtype=np.float32):
    out = np.empty(count, dtype=dtype)
    check(lib().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(dev_ptr), C.c_size_t(out.nbytes), _D2H), "hipMemcpy D2H")
    return out


def to_device(dev_ptr, a):
    a = np.ascontiguousarray(a)
    check(lib().hipMemcpy(C.c_void_p(dev_ptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), _H2D), "hipMemcpy H2D")


def copy_device(dst_ptr, src_ptr, nbytes):
    check(lib().hipMemcpy(C.c_void_p(dst_ptr), C.c_void_p(src_ptr), C.c_size_t(nbytes), _D2D), "hipMemcpy D2D")


def synchronize():
    check(lib().hipDeviceSynchronize(), "hipDeviceSynchronize")


class Stream:
    def __init__(self):
        self.handle = C.c_void_p()
        check(lib().hipStreamCreateWithFlags(C.byref(self.handle), C.c_uint(1)), "hipStreamCreateWithFlags")

    def synchronize(self):
        check(lib().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            lib().hipStreamDestroy(self.handle)
            self.handle = C.c_void_p()


class Event:
    def __init__(self):
        self.handle = C.c_void_p()
        check(lib().hipEventCreate(C.byref(self.handle)), "hipEventCreate")

    def record(self, stream=None):
        check(lib().hipEventRecord(self.handle, stream.handle if stream is not None else None), "hipEventRecord")

    def synchronize(self):
        check(lib().hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, start):
        ms = C.c_float()
        check(lib().hipEventElapsedTime(C.byref(ms), start.handle, self.handle), "hipEventElapsedTime")
        return ms.value

    def destroy(self):
        if self.handle:
            lib().hipEventDestroy(self.handle)
            self.handle = C.c_void_p()
