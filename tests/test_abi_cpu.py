"""C-ABI contract checks that need no GPU: exported symbols, struct layout, and the
argument-validation paths that return before any HIP call (reference
backend-cuda.cu:516-528, :536-545, :723, :730)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import ROOT, sf

HEADERS = [os.path.join(ROOT, "include", h) for h in ("stable-fluids-3d.h", "stable-fluids-3d-ext.h", "stable-fluids-unet.h")]


def declared_entry_points():
    names = []
    for h in HEADERS:
        text = open(h).read()
        names += re.findall(r"STABLE_FLUIDS_API\s+[\w\s\*]+?\b(stable_fluids_\w+)\s*\(", text)
    return names


def test_library_exports_every_declared_symbol():
    lib = sf.load_library()
    names = declared_entry_points()
    assert len(names) >= 7
    for n in names:
        assert hasattr(lib, n), n
    # the reference's seven names are exported unchanged
    for n in sf.runtime.ENTRY_POINTS:
        assert n in names


def test_exported_symbol_table_matches_headers():
    out = subprocess.run(["nm", "-D", "--defined-only", sf.runtime.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert exported == set(declared_entry_points())


LAYOUT_PROGRAM = r"""
#include <stddef.h>
#include <stdio.h>
#include "stable-fluids-3d-ext.h"
#define P(T) printf(#T " %zu\n", sizeof(T));
#define O(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
  P(StableFluidsFlowBoundaryFaceDesc) P(StableFluidsFlowBoundaryConfig) P(StableFluidsScalarBoundaryConfig)
  P(StableFluidsSimulationConfig) P(StableFluidsScalarFieldDesc) P(StableFluidsVectorFieldDesc)
  P(StableFluidsContextCreateDesc) P(StableFluidsViewRequest) P(StableFluidsView) P(StableFluidsExtOptions)
  P(StableFluidsExtForceAnimation) O(StableFluidsExtForceAnimation, start_step)
  O(StableFluidsContextCreateDesc, stream) O(StableFluidsContextCreateDesc, scalar_fields)
  O(StableFluidsContextCreateDesc, vector_field_count) O(StableFluidsSimulationConfig, flow_boundary)
  O(StableFluidsView, data0) O(StableFluidsViewRequest, consumer_stream) O(StableFluidsScalarFieldDesc, boundary)
  return 0;
}
"""


def test_struct_layout_matches_c_compiler(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(LAYOUT_PROGRAM)
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    a = sf.abi
    expect = {
        "StableFluidsFlowBoundaryFaceDesc": C.sizeof(a.FlowFace), "StableFluidsFlowBoundaryConfig": C.sizeof(a.FlowBoundary),
        "StableFluidsScalarBoundaryConfig": C.sizeof(a.ScalarBoundary), "StableFluidsSimulationConfig": C.sizeof(a.SimulationConfig),
        "StableFluidsScalarFieldDesc": C.sizeof(a.ScalarFieldDesc), "StableFluidsVectorFieldDesc": C.sizeof(a.VectorFieldDesc),
        "StableFluidsContextCreateDesc": C.sizeof(a.ContextCreateDesc), "StableFluidsViewRequest": C.sizeof(a.ViewRequest),
        "StableFluidsView": C.sizeof(a.View), "StableFluidsExtOptions": C.sizeof(a.ExtOptions),
        "StableFluidsExtForceAnimation": C.sizeof(a.ForceAnimation),
        "StableFluidsExtForceAnimation.start_step": a.ForceAnimation.start_step.offset,
        "StableFluidsContextCreateDesc.stream": a.ContextCreateDesc.stream.offset,
        "StableFluidsContextCreateDesc.scalar_fields": a.ContextCreateDesc.scalar_fields.offset,
        "StableFluidsContextCreateDesc.vector_field_count": a.ContextCreateDesc.vector_field_count.offset,
        "StableFluidsSimulationConfig.flow_boundary": a.SimulationConfig.flow_boundary.offset,
        "StableFluidsView.data0": a.View.data0.offset, "StableFluidsViewRequest.consumer_stream": a.ViewRequest.consumer_stream.offset,
        "StableFluidsScalarFieldDesc.boundary": a.ScalarFieldDesc.boundary.offset,
    }
    for k, v in expect.items():
        assert int(got[k]) == v, k
    # the reference's layout (stable-fluids-3d.h:44-111 on LP64)
    assert int(got["StableFluidsFlowBoundaryFaceDesc"]) == 20
    assert int(got["StableFluidsFlowBoundaryConfig"]) == 120
    assert int(got["StableFluidsScalarBoundaryConfig"]) == 48
    assert int(got["StableFluidsSimulationConfig"]) == 152
    assert int(got["StableFluidsContextCreateDesc"]) == 192
    assert int(got["StableFluidsContextCreateDesc.stream"]) == 152


def _desc(abi, flow_types=None, scalar_types=None, n=(8, 8, 8)):
    flow = abi.flow_boundary(flow_types or [abi.FLOW_FREE_SLIP] * 6)
    sb = abi.scalar_boundary(scalar_types or [abi.SCALAR_ZERO_FLUX] * 6)
    return abi.Desc(*n, 1.0 / 8, 0.1, 1e-4, 2, 2, flow, scalars=[abi.scalar_desc(1e-4, 0.0, 0.0, sb)], vectors=[abi.force_desc()])


def _create(lib, abi, d):
    h = C.c_void_p(0x1234)
    sh = (C.c_uint32 * 1)(77)
    r = lib.stable_fluids_create_context_cuda(C.byref(d.desc), C.byref(h), sh, 1, None, 0)
    return r, h, sh


@pytest.mark.parametrize("types", [
    [sf.abi.FLOW_PERIODIC, sf.abi.FLOW_FREE_SLIP] + [sf.abi.FLOW_FREE_SLIP] * 4,   # unpaired periodic
    [sf.abi.FLOW_FREE_SLIP] * 3 + [7] + [sf.abi.FLOW_FREE_SLIP] * 2,               # out-of-range type
])
def test_invalid_flow_boundary_fails_before_device_use(abi, types):
    lib = sf.load_library()
    r, h, sh = _create(lib, abi, _desc(abi, flow_types=types))
    assert r == abi.RESULT_BACKEND_FAILURE
    assert h.value is None  # *out_context cleared (reference :483)
    assert sh[0] == 77      # no handle written


def test_invalid_scalar_boundary_fails(abi):
    lib = sf.load_library()
    bad = [abi.SCALAR_PERIODIC, abi.SCALAR_ZERO_FLUX] + [abi.SCALAR_ZERO_FLUX] * 4
    r, h, _ = _create(lib, abi, _desc(abi, scalar_types=bad))
    assert r == abi.RESULT_BACKEND_FAILURE and h.value is None


@pytest.mark.parametrize("n", [(0, 8, 8), (8, -1, 8)])
def test_empty_grid_rejected(abi, n):
    lib = sf.load_library()
    r, h, _ = _create(lib, abi, _desc(abi, n=n))
    assert r == abi.RESULT_BACKEND_FAILURE and h.value is None


def test_bad_ext_options_rejected(abi):
    lib = sf.load_library()
    d = _desc(abi)
    h = C.c_void_p()
    for kw in ({"smoother": 5}, {"pressure_warm_start": 9}, {"iterate_storage": 7}, {"defect_correction_interval": 4097}):
        o = sf.ext_options(**kw)
        assert lib.stable_fluids_ext_create_context_cuda(C.byref(d.desc), C.byref(o), C.byref(h), None, 0, None, 0) == abi.RESULT_BACKEND_FAILURE
    # refinement cycles with the RBGS smoother need even extents on periodic axes
    odd = _desc(abi, flow_types=[abi.FLOW_PERIODIC] * 2 + [abi.FLOW_FREE_SLIP] * 4, n=(7, 8, 8))
    for storage in (abi.STORAGE_F16, abi.STORAGE_BF16, abi.STORAGE_F32_REFINED):
        o = sf.ext_options(iterate_storage=storage)
        assert lib.stable_fluids_ext_create_context_cuda(C.byref(odd.desc), C.byref(o), C.byref(h), None, 0, None, 0) == abi.RESULT_BACKEND_FAILURE
    o = sf.ext_options()
    o.struct_size = 12
    assert lib.stable_fluids_ext_create_context_cuda(C.byref(d.desc), C.byref(o), C.byref(h), None, 0, None, 0) == abi.RESULT_BACKEND_FAILURE


def test_null_context_calls_fail_cleanly(abi):
    lib = sf.load_library()
    view = abi.View()
    req = abi.ViewRequest(abi.VIEW_FLOW_PRESSURE, 0, 0, None)
    assert lib.stable_fluids_step_cuda(None) == abi.RESULT_BACKEND_FAILURE                       # reference :723
    assert lib.stable_fluids_get_view_cuda(None, C.byref(req), C.byref(view)) == abi.RESULT_BACKEND_FAILURE  # :730
    # the reference dereferences NULL here (:684, :694); this library reports a failure
    assert lib.stable_fluids_destroy_context_cuda(None) == abi.RESULT_BACKEND_FAILURE
    assert lib.stable_fluids_update_scalar_field_cuda(None, 1, None) == abi.RESULT_BACKEND_FAILURE
    assert lib.stable_fluids_update_scalar_field_source_cuda(None, 1, None) == abi.RESULT_BACKEND_FAILURE
    assert lib.stable_fluids_update_vector_field_cuda(None, 1, None, None, None) == abi.RESULT_BACKEND_FAILURE
    assert lib.stable_fluids_ext_set_flow_velocity_cuda(None, None, None, None) == abi.RESULT_BACKEND_FAILURE
    assert lib.stable_fluids_ext_set_force_animation_cuda(None, 1, None, None, None, None) == abi.RESULT_BACKEND_FAILURE
    assert lib.stable_fluids_ext_force_animation_step_cuda(None, None) == abi.RESULT_BACKEND_FAILURE
    assert lib.stable_fluids_ext_run_stage_cuda(None, 0, 1) == abi.RESULT_BACKEND_FAILURE
    assert lib.stable_fluids_ext_stage_bytes_cuda(None, 0) == 0
    assert lib.stable_fluids_create_context_cuda(None, None, None, 0, None, 0) == abi.RESULT_BACKEND_FAILURE


def test_interop_rejects_bad_arguments_before_device_use(abi):
    """The Vulkan hand-off entry points validate descriptors, sizes and handles on the
    host before any HIP call (no GPU here; no Vulkan loader on the GPU hosts either)."""
    lib = sf.load_library()
    slot = C.c_void_p(1)
    for args in ((-1, 4096, 0, 4096, -1), (3, 0, 0, 16, -1), (3, 4096, 0, 0, -1), (3, 4096, 4097, 1, -1),
                 (3, 4096, 4000, 200, 5), (3, 4096, 2 ** 64 - 1, 2, -1)):
        assert lib.stable_fluids_ext_interop_import_cuda(*args, C.byref(slot)) == abi.RESULT_BACKEND_FAILURE, args
        assert slot.value is None
    assert lib.stable_fluids_ext_interop_import_cuda(3, 4096, 0, 4096, -1, None) == abi.RESULT_BACKEND_FAILURE
    assert lib.stable_fluids_ext_interop_destroy_cuda(None) == abi.RESULT_BACKEND_FAILURE
    assert lib.stable_fluids_ext_interop_buffer_cuda(None) is None
    req = abi.ViewRequest(abi.VIEW_FLOW_PRESSURE, 0, 0, None)
    assert lib.stable_fluids_ext_interop_export_view_cuda(None, None, C.byref(req), 1) == abi.RESULT_BACKEND_FAILURE
