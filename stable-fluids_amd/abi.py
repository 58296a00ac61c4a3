"""ctypes mirror of include/stable-fluids-3d.h and include/stable-fluids-3d-ext.h.

The layouts follow the reference's stable-fluids-3d.h:22-154 field for field; the
sizes are asserted below against the C compiler's values (flow face 20 B, flow config
120 B, scalar config 48 B, simulation config 152 B, create desc 192 B on LP64).
"""
import ctypes as C

RESULT_OK = 0
RESULT_OUT_OF_MEMORY = 1
RESULT_BACKEND_FAILURE = 2

FLOW_NO_SLIP = 0
FLOW_FREE_SLIP = 1
FLOW_OUTFLOW = 2
FLOW_PERIODIC = 3

SCALAR_FIXED = 0
SCALAR_ZERO_FLUX = 1
SCALAR_PERIODIC = 2

VECTOR_FORCE = 0

VIEW_SCALAR_DATA = 0
VIEW_SCALAR_SOURCE = 1
VIEW_VECTOR = 2
VIEW_VELOCITY = 3
VIEW_VELOCITY_MAGNITUDE = 4
VIEW_PRESSURE = 5
VIEW_DIVERGENCE = 6
# the reference's enumerator names (stable-fluids-3d.h:120-128)
VIEW_SCALAR_FIELD_DATA = VIEW_SCALAR_DATA
VIEW_SCALAR_FIELD_SOURCE = VIEW_SCALAR_SOURCE
VIEW_VECTOR_FIELD = VIEW_VECTOR
VIEW_FLOW_VELOCITY = VIEW_VELOCITY
VIEW_FLOW_VELOCITY_MAGNITUDE = VIEW_VELOCITY_MAGNITUDE
VIEW_FLOW_PRESSURE = VIEW_PRESSURE
VIEW_FLOW_DIVERGENCE = VIEW_DIVERGENCE

LAYOUT_F32_3D = 0
LAYOUT_F32_3D_SOA3 = 1

# extension enums (stable-fluids-3d-ext.h)
SMOOTHER_RBGS = 0
SMOOTHER_JACOBI = 1
STORAGE_F32 = 0
STORAGE_F16 = 1
STORAGE_BF16 = 2
STORAGE_F32_REFINED = 3
WARM_START_ZERO = 0
WARM_START_LEARNED = 1
STAGE_PRESSURE_PASS = 0
STAGE_VELOCITY_PASS = 1
STAGE_ADVECT = 2
STAGE_PROJECT = 3
STAGE_JACOBI_PASS = 4
STAGE_COPY = 5
STAGE_FULL_PRESSURE_PASS = 8
STAGE_FULL_VELOCITY_PASS = 9
STAGE_PRESSURE_SOLVE = 10
STAGE_PROJECTION = 11
TILE_128X16 = 0
TILE_64X32 = 1
TILE_64X16 = 2
TILE_32X32 = 3
TILE_32X16 = 4


class FlowFace(C.Structure):
    _fields_ = [("type", C.c_uint32), ("velocity_x", C.c_float), ("velocity_y", C.c_float),
                ("velocity_z", C.c_float), ("pressure", C.c_float)]


class FlowBoundary(C.Structure):
    _fields_ = [(n, FlowFace) for n in ("x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus")]


class ScalarFace(C.Structure):
    _fields_ = [("type", C.c_uint32), ("value", C.c_float)]


class ScalarBoundary(C.Structure):
    _fields_ = [(n, ScalarFace) for n in ("x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus")]


class SimulationConfig(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("cell_size", C.c_float),
                ("dt", C.c_float), ("viscosity", C.c_float), ("diffuse_iterations", C.c_int32),
                ("pressure_iterations", C.c_int32), ("flow_boundary", FlowBoundary)]


class ScalarFieldDesc(C.Structure):
    _fields_ = [("diffusion", C.c_float), ("dissipation", C.c_float), ("initial_value", C.c_float),
                ("boundary", ScalarBoundary)]


class VectorFieldDesc(C.Structure):
    _fields_ = [("usage", C.c_uint32), ("initial_value_x", C.c_float), ("initial_value_y", C.c_float),
                ("initial_value_z", C.c_float)]


class ContextCreateDesc(C.Structure):
    _fields_ = [("config", SimulationConfig), ("stream", C.c_void_p),
                ("scalar_fields", C.POINTER(ScalarFieldDesc)), ("scalar_field_count", C.c_uint32),
                ("vector_fields", C.POINTER(VectorFieldDesc)), ("vector_field_count", C.c_uint32)]


class ViewRequest(C.Structure):
    _fields_ = [("kind", C.c_uint32), ("scalar_field", C.c_uint32), ("vector_field", C.c_uint32),
                ("consumer_stream", C.c_void_p)]


class View(C.Structure):
    _fields_ = [("layout", C.c_uint32), ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("row_stride_bytes", C.c_uint64), ("slice_stride_bytes", C.c_uint64),
                ("data0", C.c_void_p), ("data1", C.c_void_p), ("data2", C.c_void_p)]


class ExtOptions(C.Structure):
    """StableFluidsExtOptions (stable-fluids-3d-ext.h)."""
    _fields_ = [("struct_size", C.c_uint32), ("smoother", C.c_uint32), ("iterate_storage", C.c_uint32),
                ("defect_correction_interval", C.c_uint32), ("pressure_warm_start", C.c_uint32),
                ("warm_start_iterations", C.c_uint32), ("warm_start_coarsening", C.c_uint32),
                ("flags", C.c_uint32)]


class ForceAnimation(C.Structure):
    """StableFluidsExtForceAnimation (stable-fluids-3d-ext.h)."""
    _fields_ = [("phase_rate", C.c_float), ("drift_x_amplitude", C.c_float), ("drift_z_amplitude", C.c_float),
                ("drift_z_rate", C.c_float), ("swirl_amplitude", C.c_float), ("swirl_rate", C.c_float),
                ("start_step", C.c_uint32)]


def plume_scene_animation(start_step=0):
    """The swirl/drift schedule of the reference's plume scene (vulkan-app/scene_plume.cpp:327-330)."""
    return ForceAnimation(0.045, 1.15, 0.85, 0.71, 1.65, 0.53, start_step)


assert C.sizeof(FlowFace) == 20 and C.sizeof(FlowBoundary) == 120
assert C.sizeof(ScalarBoundary) == 48 and C.sizeof(SimulationConfig) == 152
assert C.sizeof(ContextCreateDesc) == 192 and ContextCreateDesc.stream.offset == 152
assert C.sizeof(ExtOptions) == 32 and C.sizeof(ForceAnimation) == 28


def flow_boundary(types, velocity=None, pressure=None):
    """types: 6 face types in the order x-, x+, y-, y+, z-, z+."""
    fb = FlowBoundary()
    names = ("x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus")
    for i, n in enumerate(names):
        f = getattr(fb, n)
        f.type = types[i]
        if velocity is not None:
            f.velocity_x, f.velocity_y, f.velocity_z = velocity[i]
        if pressure is not None:
            f.pressure = pressure[i]
    return fb


def scalar_boundary(types, values=None):
    sb = ScalarBoundary()
    names = ("x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus")
    for i, n in enumerate(names):
        f = getattr(sb, n)
        f.type = types[i]
        if values is not None:
            f.value = values[i]
    return sb


class Desc:
    """Owns a ContextCreateDesc together with the arrays it points into."""

    def __init__(self, nx, ny, nz, cell_size, dt, viscosity, diffuse_iterations, pressure_iterations,
                 flow, scalars=(), vectors=(), stream=None):
        self.scalar_array = (ScalarFieldDesc * max(1, len(scalars)))(*scalars)
        self.vector_array = (VectorFieldDesc * max(1, len(vectors)))(*vectors)
        cfg = SimulationConfig(nx, ny, nz, cell_size, dt, viscosity, diffuse_iterations, pressure_iterations, flow)
        self.desc = ContextCreateDesc(cfg, stream, self.scalar_array, len(scalars), self.vector_array, len(vectors))
        self.shape = (nz, ny, nx)
        self.n_scalars = len(scalars)
        self.n_vectors = len(vectors)

    @property
    def cells(self):
        return self.shape[0] * self.shape[1] * self.shape[2]


def scalar_desc(diffusion, dissipation=0.0, initial_value=0.0, boundary=None):
    if boundary is None:
        boundary = scalar_boundary([SCALAR_ZERO_FLUX] * 6)
    return ScalarFieldDesc(diffusion, dissipation, initial_value, boundary)


def force_desc(ix=0.0, iy=0.0, iz=0.0):
    return VectorFieldDesc(VECTOR_FORCE, ix, iy, iz)
