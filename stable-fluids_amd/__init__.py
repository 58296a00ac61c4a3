"""stable-fluids on MI355X (gfx950): Python mirror of the C ABI in libstable-fluids.so.

The product is the HIP library built from csrc/ (C ABI in include/stable-fluids-3d.h and
include/stable-fluids-3d-ext.h). This package only binds it for Python callers, tests and
the benchmark: `runtime.Context` mirrors the reference's entry points, `scenes` builds
the workloads, `hip` holds device-buffer plumbing.
"""
from . import abi, hip, runtime, scenes, unet, warm_eval  # noqa: F401
from .runtime import Batch, Context, StableFluidsError, ext_options, load_library  # noqa: F401
