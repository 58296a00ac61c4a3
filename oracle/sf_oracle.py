"""ctypes wrapper of the host-loop oracle (libsf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg to check or time against; the product library never uses it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsf_oracle.so")
RBGS, JACOBI = 0, 1
_lib = None


def build():
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def load(abi):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp, u32, i32, d = C.c_void_p, C.c_uint32, C.c_int, C.c_double
    dp = C.POINTER(C.c_double)
    fb, sb = C.POINTER(abi.FlowBoundary), C.POINTER(abi.ScalarBoundary)
    lib.sfo_create.argtypes = [C.POINTER(abi.ContextCreateDesc), i32, i32]
    lib.sfo_create.restype = vp
    lib.sfo_destroy.argtypes = [vp]
    lib.sfo_update_scalar_field.argtypes = [vp, u32, vp]
    lib.sfo_update_scalar_field_source.argtypes = [vp, u32, vp]
    lib.sfo_update_vector_field.argtypes = [vp, u32, vp, vp, vp]
    lib.sfo_set_flow_velocity.argtypes = [vp, vp, vp, vp]
    lib.sfo_step.argtypes = [vp, i32]
    lib.sfo_get_field.argtypes = [vp, u32, u32, dp, dp, dp]
    lib.sfo_load_scalar.argtypes = [dp, i32, i32, i32, i32, i32, i32, sb]
    lib.sfo_load_scalar.restype = d
    lib.sfo_load_pressure.argtypes = [dp, i32, i32, i32, i32, i32, i32, fb]
    lib.sfo_load_pressure.restype = d
    lib.sfo_load_velocity.argtypes = [dp, i32, i32, i32, i32, i32, i32, i32, fb]
    lib.sfo_load_velocity.restype = d
    lib.sfo_sample_scalar.argtypes = [dp, d, d, d, i32, i32, i32, d, sb]
    lib.sfo_sample_scalar.restype = d
    lib.sfo_sample_velocity.argtypes = [dp, i32, d, d, d, i32, i32, i32, d, fb]
    lib.sfo_sample_velocity.restype = d
    lib.sfo_trace_rk2.argtypes = [dp, dp, dp, d, d, d, d, i32, i32, i32, d, fb, dp]
    lib.sfo_advect_velocity.argtypes = [dp, dp, i32, dp, dp, dp, d, i32, i32, i32, d, fb]
    lib.sfo_advect_scalar.argtypes = [dp, dp, dp, dp, dp, d, i32, i32, i32, d, sb, fb]
    lib.sfo_smooth_scalar.argtypes = [dp, dp, d, i32, i32, i32, i32, i32, sb]
    lib.sfo_smooth_velocity.argtypes = [dp, dp, d, i32, i32, i32, i32, i32, i32, fb]
    lib.sfo_smooth_pressure.argtypes = [dp, dp, d, i32, i32, i32, i32, i32, fb]
    lib.sfo_divergence.argtypes = [dp, dp, dp, dp, i32, i32, i32, d, fb]
    lib.sfo_project.argtypes = [dp, dp, dp, dp, i32, i32, i32, d, fb]
    lib.sfo_threads.restype = i32
    lib.sfo_set_threads.argtypes = [i32]
    _lib = lib
    return lib


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


def fptr(a):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.ctypes.data_as(C.c_void_p), a


class OracleSim:
    """The whole-step oracle with the same calls as the C ABI (host arrays)."""

    def __init__(self, abi, desc, precision=64, smoother=RBGS):
        self.lib = load(abi)
        self.desc = desc
        self.shape = desc.shape
        self.handle = self.lib.sfo_create(C.byref(desc.desc), precision, smoother)
        if not self.handle:
            raise RuntimeError("sfo_create failed")

    def _f(self, a):
        if a is None:
            return None, None
        a = np.ascontiguousarray(a, dtype=np.float32).ravel()
        return a.ctypes.data_as(C.c_void_p), a

    def update_scalar_field(self, h, values):
        p, keep = self._f(values)
        return self.lib.sfo_update_scalar_field(self.handle, h, p)

    def update_scalar_field_source(self, h, values):
        p, keep = self._f(values)
        return self.lib.sfo_update_scalar_field_source(self.handle, h, p)

    def update_vector_field(self, h, x, y, z):
        (px, kx), (py, ky), (pz, kz) = self._f(x), self._f(y), self._f(z)
        return self.lib.sfo_update_vector_field(self.handle, h, px, py, pz)

    def set_flow_velocity(self, x, y, z):
        (px, kx), (py, ky), (pz, kz) = self._f(x), self._f(y), self._f(z)
        return self.lib.sfo_set_flow_velocity(self.handle, px, py, pz)

    def step(self, steps=1):
        self.lib.sfo_step(self.handle, steps)

    def read(self, kind, handle=0):
        n = self.shape[0] * self.shape[1] * self.shape[2]
        outs = [np.empty(n, np.float64) for _ in range(3)]
        r = self.lib.sfo_get_field(self.handle, kind, handle, dptr(outs[0]), dptr(outs[1]), dptr(outs[2]))
        if r != 0:
            raise RuntimeError(f"sfo_get_field returned {r}")
        if kind in (2, 3):
            return [o.reshape(self.shape) for o in outs]
        return outs[0].reshape(self.shape)

    def destroy(self):
        if self.handle:
            self.lib.sfo_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def threads(abi):
    return load(abi).sfo_threads()


def set_threads(abi, n):
    load(abi).sfo_set_threads(int(n))
