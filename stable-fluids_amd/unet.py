"""Python mirror of include/stable-fluids-unet.h (the pressure warm-start network).

Parameters travel as host float32 arrays in the public layout (11 layers, each weights
[27][cout][cin] then bias [cout]); tensors are device buffers (hip.DeviceBuffer or raw
device pointers). Every call runs libstable-fluids.so's HIP kernels; there is no host
fallback.
"""
import ctypes as C

import numpy as np

from . import abi, hip
from .runtime import StableFluidsError, _ptr, load_library

LOSS_CENTERED = 1  # STABLE_FLUIDS_UNET_LOSS_CENTERED


def param_count():
    return int(load_library().stable_fluids_unet_param_count())


def _check(name, r):
    if r != abi.RESULT_OK:
        raise StableFluidsError(name, r)


class UNet:
    """Network for an nx*ny*nz grid (extents multiples of 4) on `stream` (None: own stream)."""

    def __init__(self, nx, ny, nz, stream=None):
        self.lib = load_library()
        self.shape = (nz, ny, nx)
        h = C.c_void_p()
        s = stream.handle if isinstance(stream, hip.Stream) else stream
        _check("stable_fluids_unet_create_cuda", self.lib.stable_fluids_unet_create_cuda(nx, ny, nz, s, C.byref(h)))
        self.handle = h.value
        self.count = param_count()

    def destroy(self):
        if self.handle:
            self.lib.stable_fluids_unet_destroy_cuda(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == self.count
        _check("set_params", self.lib.stable_fluids_unet_set_params_cuda(self.handle, p.ctypes.data_as(C.c_void_p), self.count))

    def _get(self, name):
        out = np.empty(self.count, dtype=np.float32)
        _check(name, getattr(self.lib, f"stable_fluids_unet_{name}_cuda")(self.handle, out.ctypes.data_as(C.c_void_p), self.count))
        return out

    def params(self):
        return self._get("get_params")

    def grads(self):
        return self._get("get_grads")

    def forward(self, inp, out):
        """inp: device [v][2] float32; out: device [v] float32 (asynchronous)."""
        _check("forward", self.lib.stable_fluids_unet_forward_cuda(self.handle, _ptr(inp), _ptr(out)))

    def backward(self, inp, target, loss=None, weight=1.0, centered=False):
        flags = LOSS_CENTERED if centered else 0
        _check("backward", self.lib.stable_fluids_unet_backward_cuda(self.handle, _ptr(inp), _ptr(target), weight, flags, _ptr(loss)))

    def adam_step(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        _check("adam_step", self.lib.stable_fluids_unet_adam_step_cuda(self.handle, lr, beta1, beta2, eps))


def conv3(x, w, bias, y, cin, cout, shape, relu=False, stream=None):
    nz, ny, nx = shape
    _check("conv3", load_library().stable_fluids_unet_conv3_cuda(stream, cin, cout, _ptr(x), _ptr(w), _ptr(bias), _ptr(y), int(relu), nx, ny, nz))


def conv3_wgrad(x, g, dw, db, cin, cout, shape, stream=None):
    nz, ny, nx = shape
    _check("conv3_wgrad", load_library().stable_fluids_unet_conv3_wgrad_cuda(stream, cin, cout, _ptr(x), _ptr(g), _ptr(dw), _ptr(db), nx, ny, nz))


def to_bf16_bits(a):
    """float -> BF16 bit patterns (round to nearest even), uint16."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def from_bf16_bits(h):
    return (np.asarray(h, dtype=np.uint32) << 16).view(np.float32)
