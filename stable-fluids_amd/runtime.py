"""Host-side mirror of the stable-fluids C ABI over libstable-fluids.so.

`Context` wraps the seven entry points of include/stable-fluids-3d.h (reference
stable-fluids-3d.h:112-154) with the same names, argument meaning and error behaviour:
every call returns the library's StableFluidsResult and `check=True` turns a non-zero
result into `StableFluidsError`. Device buffers come from hip.DeviceBuffer. There is no
CPU fallback: loading fails loudly when the HIP library has not been built.
"""
import ctypes as C
import os

import numpy as np

from . import abi, hip

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libstable-fluids.so")
_lib = None

ENTRY_POINTS = (
    "stable_fluids_create_context_cuda", "stable_fluids_destroy_context_cuda",
    "stable_fluids_update_scalar_field_cuda", "stable_fluids_update_scalar_field_source_cuda",
    "stable_fluids_update_vector_field_cuda", "stable_fluids_step_cuda", "stable_fluids_get_view_cuda",
)
EXT_ENTRY_POINTS = (
    "stable_fluids_ext_create_context_cuda", "stable_fluids_ext_set_flow_velocity_cuda",
    "stable_fluids_ext_set_flow_pressure_cuda", "stable_fluids_ext_field_norms_cuda",
    "stable_fluids_ext_run_stage_cuda", "stable_fluids_ext_stage_bytes_cuda", "stable_fluids_ext_get_stream_cuda",
    "stable_fluids_ext_set_pressure_model_cuda", "stable_fluids_ext_warm_start_stats_cuda", "stable_fluids_ext_training_pair_cuda",
    "stable_fluids_ext_get_pressure_rhs_cuda",
    "stable_fluids_ext_batch_create_cuda", "stable_fluids_ext_batch_destroy_cuda", "stable_fluids_ext_batch_step_cuda",
    "stable_fluids_ext_batch_get_stream_cuda", "stable_fluids_ext_batch_set_z_chunk_cuda", "stable_fluids_ext_batch_info_cuda",
    "stable_fluids_ext_batch_autotune_cuda", "stable_fluids_ext_set_force_animation_cuda",
    "stable_fluids_ext_force_animation_step_cuda", "stable_fluids_ext_field_norms_noalloc_cuda",
    "stable_fluids_ext_set_smoother_tiling_cuda", "stable_fluids_ext_get_smoother_tiling_cuda",
    "stable_fluids_ext_autotune_cuda", "stable_fluids_ext_batch_set_tiling_cuda", "stable_fluids_ext_batch_get_tile_cuda",
    "stable_fluids_ext_interop_import_cuda", "stable_fluids_ext_interop_destroy_cuda", "stable_fluids_ext_interop_buffer_cuda",
    "stable_fluids_ext_interop_export_view_cuda",
)
UNET_ENTRY_POINTS = (
    "stable_fluids_unet_param_count", "stable_fluids_unet_create_cuda", "stable_fluids_unet_destroy_cuda",
    "stable_fluids_unet_set_params_cuda", "stable_fluids_unet_get_params_cuda", "stable_fluids_unet_get_grads_cuda",
    "stable_fluids_unet_forward_cuda", "stable_fluids_unet_backward_cuda", "stable_fluids_unet_adam_step_cuda",
    "stable_fluids_unet_conv3_cuda", "stable_fluids_unet_conv3_wgrad_cuda",
)


class StableFluidsError(RuntimeError):
    def __init__(self, what, code):
        super().__init__(f"{what} returned {code}")
        self.code = code


class ExtNorms(C.Structure):
    _fields_ = [("divergence_l2", C.c_double), ("divergence_max", C.c_double),
                ("velocity_l2", C.c_double), ("pressure_l2", C.c_double)]


def load_library(path=LIB_PATH):
    """Loads libstable-fluids.so and declares every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    vp, u32, fp = C.c_void_p, C.c_uint32, C.POINTER(C.c_float)
    lib.stable_fluids_create_context_cuda.argtypes = [C.POINTER(abi.ContextCreateDesc), C.POINTER(vp), C.POINTER(u32), u32, C.POINTER(u32), u32]
    lib.stable_fluids_ext_create_context_cuda.argtypes = [C.POINTER(abi.ContextCreateDesc), C.POINTER(abi.ExtOptions), C.POINTER(vp), C.POINTER(u32), u32, C.POINTER(u32), u32]
    lib.stable_fluids_destroy_context_cuda.argtypes = [vp]
    lib.stable_fluids_update_scalar_field_cuda.argtypes = [vp, u32, vp]
    lib.stable_fluids_update_scalar_field_source_cuda.argtypes = [vp, u32, vp]
    lib.stable_fluids_update_vector_field_cuda.argtypes = [vp, u32, vp, vp, vp]
    lib.stable_fluids_step_cuda.argtypes = [vp]
    lib.stable_fluids_get_view_cuda.argtypes = [vp, C.POINTER(abi.ViewRequest), C.POINTER(abi.View)]
    lib.stable_fluids_ext_set_flow_velocity_cuda.argtypes = [vp, vp, vp, vp]
    lib.stable_fluids_ext_set_force_animation_cuda.argtypes = [vp, u32, C.POINTER(abi.ForceAnimation), vp, vp, vp]
    lib.stable_fluids_ext_force_animation_step_cuda.argtypes = [vp, C.POINTER(u32)]
    lib.stable_fluids_ext_set_flow_pressure_cuda.argtypes = [vp, vp]
    lib.stable_fluids_ext_field_norms_cuda.argtypes = [vp, C.POINTER(ExtNorms)]
    lib.stable_fluids_ext_field_norms_noalloc_cuda.argtypes = [vp, vp, vp, vp, vp, C.c_uint64, vp, vp]
    lib.stable_fluids_ext_set_smoother_tiling_cuda.argtypes = [vp, C.c_uint32, C.c_uint32]
    lib.stable_fluids_ext_get_smoother_tiling_cuda.argtypes = [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.stable_fluids_ext_autotune_cuda.argtypes = [vp, C.c_char_p, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.stable_fluids_ext_interop_import_cuda.argtypes = [C.c_int, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int, C.POINTER(vp)]
    lib.stable_fluids_ext_interop_destroy_cuda.argtypes = [vp]
    lib.stable_fluids_ext_interop_buffer_cuda.argtypes = [vp]
    lib.stable_fluids_ext_interop_buffer_cuda.restype = vp
    lib.stable_fluids_ext_interop_export_view_cuda.argtypes = [vp, vp, C.POINTER(abi.ViewRequest), C.c_uint64]
    lib.stable_fluids_ext_run_stage_cuda.argtypes = [vp, u32, u32]
    lib.stable_fluids_ext_stage_bytes_cuda.argtypes = [vp, u32]
    lib.stable_fluids_ext_stage_bytes_cuda.restype = C.c_uint64
    lib.stable_fluids_ext_get_stream_cuda.argtypes = [vp]
    lib.stable_fluids_ext_get_stream_cuda.restype = vp
    i32, u64, f32 = C.c_int32, C.c_uint64, C.c_float
    lib.stable_fluids_ext_set_pressure_model_cuda.argtypes = [vp, vp, u64]
    lib.stable_fluids_ext_batch_create_cuda.argtypes = [C.POINTER(vp), u32, vp, C.POINTER(vp)]
    lib.stable_fluids_ext_batch_destroy_cuda.argtypes = [vp]
    lib.stable_fluids_ext_batch_step_cuda.argtypes = [vp]
    lib.stable_fluids_ext_batch_get_stream_cuda.argtypes = [vp]
    lib.stable_fluids_ext_batch_get_stream_cuda.restype = vp
    lib.stable_fluids_ext_batch_set_z_chunk_cuda.argtypes = [vp, u32, u32]
    lib.stable_fluids_ext_batch_set_tiling_cuda.argtypes = [vp, u32, u32, u32]
    lib.stable_fluids_ext_batch_get_tile_cuda.argtypes = [vp, u32, C.POINTER(u32)]
    lib.stable_fluids_ext_batch_info_cuda.argtypes = [vp, u32, C.POINTER(C.c_int32), C.POINTER(u32)]
    lib.stable_fluids_ext_batch_autotune_cuda.argtypes = [vp, C.c_char_p, u32, C.POINTER(u32)]
    lib.stable_fluids_ext_get_pressure_rhs_cuda.argtypes = [vp]
    lib.stable_fluids_ext_get_pressure_rhs_cuda.restype = vp
    lib.stable_fluids_ext_warm_start_stats_cuda.argtypes = [vp, C.POINTER(C.c_double)]
    lib.stable_fluids_ext_training_pair_cuda.argtypes = [vp, vp, vp, u32, C.c_double, u32, C.POINTER(C.c_double), C.POINTER(u32)]
    lib.stable_fluids_unet_param_count.argtypes = []
    lib.stable_fluids_unet_param_count.restype = u64
    lib.stable_fluids_unet_create_cuda.argtypes = [i32, i32, i32, vp, C.POINTER(vp)]
    lib.stable_fluids_unet_destroy_cuda.argtypes = [vp]
    for name in ("set_params", "get_params", "get_grads"):
        getattr(lib, f"stable_fluids_unet_{name}_cuda").argtypes = [vp, vp, u64]
    lib.stable_fluids_unet_forward_cuda.argtypes = [vp, vp, vp]
    lib.stable_fluids_unet_backward_cuda.argtypes = [vp, vp, vp, f32, u32, vp]
    lib.stable_fluids_unet_adam_step_cuda.argtypes = [vp, f32, f32, f32, f32]
    lib.stable_fluids_unet_conv3_cuda.argtypes = [vp, i32, i32, vp, vp, vp, vp, i32, i32, i32, i32]
    lib.stable_fluids_unet_conv3_wgrad_cuda.argtypes = [vp, i32, i32, vp, vp, vp, vp, i32, i32, i32]
    for name in ENTRY_POINTS + EXT_ENTRY_POINTS + UNET_ENTRY_POINTS:
        if name not in ("stable_fluids_ext_stage_bytes_cuda", "stable_fluids_ext_get_stream_cuda", "stable_fluids_unet_param_count",
                        "stable_fluids_ext_get_pressure_rhs_cuda", "stable_fluids_ext_batch_get_stream_cuda",
                        "stable_fluids_ext_interop_buffer_cuda"):
            getattr(lib, name).restype = C.c_int
    _lib = lib
    return lib


def ext_options(smoother=abi.SMOOTHER_RBGS, iterate_storage=abi.STORAGE_F32, defect_correction_interval=0,
                pressure_warm_start=abi.WARM_START_ZERO, warm_start_iterations=0, warm_start_coarsening=0):
    o = abi.ExtOptions()
    o.struct_size = C.sizeof(abi.ExtOptions)
    o.smoother = smoother
    o.iterate_storage = iterate_storage
    o.defect_correction_interval = defect_correction_interval
    o.pressure_warm_start = pressure_warm_start
    o.warm_start_iterations = warm_start_iterations
    o.warm_start_coarsening = warm_start_coarsening
    o.flags = 0
    return o


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, hip.DeviceBuffer):
        return x.value
    return int(x)


class Context:
    """One simulation context (reference ContextStorage, backend-cuda.cu:11-56)."""

    def __init__(self, desc, options=None, stream=None):
        lib = load_library()
        self.lib = lib
        self.desc = desc
        # the desc may be reused: always write this context's stream (NULL: the library's own)
        desc.desc.stream = (stream.handle if isinstance(stream, hip.Stream) else stream) if stream is not None else None
        self.scalar_handles = (C.c_uint32 * max(1, desc.n_scalars))()
        self.vector_handles = (C.c_uint32 * max(1, desc.n_vectors))()
        handle = C.c_void_p()
        if options is None:
            r = lib.stable_fluids_create_context_cuda(C.byref(desc.desc), C.byref(handle), self.scalar_handles, desc.n_scalars, self.vector_handles, desc.n_vectors)
        else:
            r = lib.stable_fluids_ext_create_context_cuda(C.byref(desc.desc), C.byref(options), C.byref(handle), self.scalar_handles, desc.n_scalars, self.vector_handles, desc.n_vectors)
        if r != abi.RESULT_OK:
            raise StableFluidsError("stable_fluids_create_context_cuda", r)
        self.handle = handle
        self.shape = desc.shape
        self.cells = desc.cells
        self._staging = []

    # ---- reference entry points ----
    def destroy(self):
        if self.handle:
            r = self.lib.stable_fluids_destroy_context_cuda(self.handle)
            self.handle = C.c_void_p()
            self._staging.clear()
            return r
        return abi.RESULT_OK

    def update_scalar_field(self, handle, values):
        return self._call("stable_fluids_update_scalar_field_cuda", handle, _ptr(values))

    def update_scalar_field_source(self, handle, values):
        return self._call("stable_fluids_update_scalar_field_source_cuda", handle, _ptr(values))

    def update_vector_field(self, handle, x, y, z):
        return self._call("stable_fluids_update_vector_field_cuda", handle, _ptr(x), _ptr(y), _ptr(z))

    def step(self, steps=1):
        for _ in range(steps):
            r = self.lib.stable_fluids_step_cuda(self.handle)
            if r != abi.RESULT_OK:
                raise StableFluidsError("stable_fluids_step_cuda", r)

    def get_view(self, kind, scalar_field=0, vector_field=0, consumer_stream=None):
        req = abi.ViewRequest(kind, scalar_field, vector_field, consumer_stream)
        view = abi.View()
        r = self.lib.stable_fluids_get_view_cuda(self.handle, C.byref(req), C.byref(view))
        if r != abi.RESULT_OK:
            raise StableFluidsError("stable_fluids_get_view_cuda", r)
        return view

    # ---- extensions ----
    def set_force_animation(self, handle, animation, drift=None, swirl_x=None, swirl_z=None):
        """Animated force on field `handle` (abi.ForceAnimation, or None to stop); the masks
        are device buffers (see upload)."""
        a = C.byref(animation) if animation is not None else None
        return self._call("stable_fluids_ext_set_force_animation_cuda", handle, a, _ptr(drift), _ptr(swirl_x), _ptr(swirl_z))

    def force_animation_step(self):
        out = C.c_uint32(0)
        self._call("stable_fluids_ext_force_animation_step_cuda", C.byref(out))
        return out.value

    def set_flow_velocity(self, x, y, z):
        return self._call("stable_fluids_ext_set_flow_velocity_cuda", _ptr(x), _ptr(y), _ptr(z))

    def set_flow_pressure(self, p):
        return self._call("stable_fluids_ext_set_flow_pressure_cuda", _ptr(p))

    def norms(self):
        out = ExtNorms()
        r = self.lib.stable_fluids_ext_field_norms_cuda(self.handle, C.byref(out))
        if r != abi.RESULT_OK:
            raise StableFluidsError("stable_fluids_ext_field_norms_cuda", r)
        return {"divergence_l2": out.divergence_l2, "divergence_max": out.divergence_max,
                "velocity_l2": out.velocity_l2, "pressure_l2": out.pressure_l2}

    def set_smoother_tiling(self, tile, z_chunk=0):
        return self._call("stable_fluids_ext_set_smoother_tiling_cuda", tile, z_chunk)

    def smoother_tiling(self):
        t, z = C.c_uint32(0), C.c_uint32(0)
        self._call("stable_fluids_ext_get_smoother_tiling_cuda", C.byref(t), C.byref(z))
        return t.value, z.value

    def autotune(self, cache_path=None, repeats=10):
        """Times every smoother tiling and keeps the fastest; returns (tile, z_chunk)."""
        t, z = C.c_uint32(0), C.c_uint32(0)
        self._call("stable_fluids_ext_autotune_cuda", cache_path.encode() if cache_path else None, repeats, C.byref(t), C.byref(z))
        return t.value, z.value

    def run_stage(self, stage, repeat):
        return self._call("stable_fluids_ext_run_stage_cuda", stage, repeat)

    def time_stage(self, stage, repeat, warmup=2):
        """Mean GPU milliseconds of one run of `stage`: HIP events recorded on the context
        stream around `repeat` runs, after `warmup` untimed ones."""
        s = C.c_void_p(self.stream)
        ev0, ev1 = hip.Event(), hip.Event()
        try:
            if warmup:
                self.run_stage(stage, warmup)
            hip.check(hip.lib().hipEventRecord(ev0.handle, s), "hipEventRecord")
            self.run_stage(stage, repeat)
            hip.check(hip.lib().hipEventRecord(ev1.handle, s), "hipEventRecord")
            ev1.synchronize()
            return ev1.elapsed_ms(ev0) / repeat
        finally:
            ev0.destroy()
            ev1.destroy()

    def set_pressure_model(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32)
        return self._call("stable_fluids_ext_set_pressure_model_cuda", p.ctypes.data_as(C.c_void_p), p.size)

    def warm_start_stats(self):
        out = (C.c_double * 3)()
        self._call("stable_fluids_ext_warm_start_stats_cuda", out)
        return {"guess_residual": out[0], "zero_residual": out[1], "fallback": bool(out[2])}

    def pressure_rhs(self):
        """Host copy of the last step's pre-projection divergence, [nz][ny][nx] float32."""
        hip.synchronize()
        ptr = self.lib.stable_fluids_ext_get_pressure_rhs_cuda(self.handle)
        return hip.to_host(ptr, self.desc.cells).reshape(self.desc.shape)

    def training_pair(self, inp, target, coarsening=0, rel_tol=1e-6, max_iterations=20000):
        """Fills device buffers inp [v][2] and target [v] on the grid coarsened 2^coarsening
        times; returns (relative residual, iterations)."""
        res, its = C.c_double(), C.c_uint32()
        self._call("stable_fluids_ext_training_pair_cuda", _ptr(inp), _ptr(target), coarsening, rel_tol, max_iterations, C.byref(res), C.byref(its))
        return res.value, its.value

    def stage_bytes(self, stage):
        return int(self.lib.stable_fluids_ext_stage_bytes_cuda(self.handle, stage))

    @property
    def stream(self):
        return self.lib.stable_fluids_ext_get_stream_cuda(self.handle)

    # ---- host conveniences (upload from numpy, read views back) ----
    def upload(self, array):
        """Copies a host array into a new device buffer kept alive with the context."""
        buf = hip.DeviceBuffer.from_array(np.ascontiguousarray(array, dtype=np.float32).ravel())
        self._staging.append(buf)
        return buf

    def synchronize(self):
        hip.check(hip.lib().hipStreamSynchronize(C.c_void_p(self.stream)), "hipStreamSynchronize")

    def read(self, kind, handle=0):
        """Returns the view's field(s) as numpy arrays shaped (nz, ny, nx)."""
        self.synchronize()
        view = self.get_view(kind, handle if kind in (abi.VIEW_SCALAR_DATA, abi.VIEW_SCALAR_SOURCE) else 0,
                             handle if kind == abi.VIEW_VECTOR else 0)
        ptrs = [view.data0] + ([view.data1, view.data2] if view.layout == abi.LAYOUT_F32_3D_SOA3 else [])
        out = [hip.to_host(p, self.cells).reshape(self.shape) for p in ptrs]
        return out[0] if len(out) == 1 else out

    def _call(self, name, *args):
        r = getattr(self.lib, name)(self.handle, *args)
        if r != abi.RESULT_OK:
            raise StableFluidsError(name, r)
        return r

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


class Batch:
    """Many contexts stepped together, one kernel launch per stage
    (stable_fluids_ext_batch_*_cuda). The contexts stay owned by the caller."""

    def __init__(self, contexts, stream=None):
        self.lib = load_library()
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.handle for c in self.contexts])
        h = C.c_void_p()
        s = stream.handle if isinstance(stream, hip.Stream) else stream
        r = self.lib.stable_fluids_ext_batch_create_cuda(arr, len(self.contexts), s, C.byref(h))
        if r != abi.RESULT_OK:
            raise StableFluidsError("stable_fluids_ext_batch_create_cuda", r)
        self.handle = h.value

    def _call(self, name, *args):
        r = getattr(self.lib, name)(self.handle, *args)
        if r != abi.RESULT_OK:
            raise StableFluidsError(name, r)
        return r

    def step(self, n=1):
        for _ in range(n):
            self._call("stable_fluids_ext_batch_step_cuda")

    @property
    def stream(self):
        return self.lib.stable_fluids_ext_batch_get_stream_cuda(self.handle)

    def set_z_chunk(self, sim, z_chunk):
        self._call("stable_fluids_ext_batch_set_z_chunk_cuda", sim, z_chunk)

    def set_tiling(self, sim, tile, z_chunk=0):
        self._call("stable_fluids_ext_batch_set_tiling_cuda", sim, tile, z_chunk)

    def info(self, sim):
        t = (C.c_int32 * 4)()
        n, tile = C.c_uint32(), C.c_uint32()
        self._call("stable_fluids_ext_batch_info_cuda", sim, t, C.byref(n))
        self._call("stable_fluids_ext_batch_get_tile_cuda", sim, C.byref(tile))
        return {"tiles_x": t[0], "tiles_y": t[1], "chunks": t[2], "z_chunk": t[3], "tile": tile.value,
                "launches_per_step": n.value}

    def autotune(self, cache_path=None, repeats=10):
        n = C.c_uint32()
        self._call("stable_fluids_ext_batch_autotune_cuda", cache_path.encode() if cache_path else None, repeats, C.byref(n))
        return n.value

    def destroy(self):
        if self.handle:
            self.lib.stable_fluids_ext_batch_destroy_cuda(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()
