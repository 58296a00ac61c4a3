"""C-ABI behaviour on the GPU: handles, NULL updates, views, stream hand-off, extension
entry points (reference backend-cuda.cu:692-800)."""
import ctypes as C

import numpy as np
import pytest

from conftest import sf

abi = sf.abi
pytestmark = pytest.mark.gpu


def small_desc(nx=10, ny=8, nz=6, scalars=2, vectors=1):
    flow = abi.flow_boundary([abi.FLOW_FREE_SLIP] * 6)
    s = [abi.scalar_desc(1e-3, 0.0, 0.5 + i) for i in range(scalars)]
    v = [abi.force_desc(0.1 * (i + 1), 0.2, 0.3) for i in range(vectors)]
    return abi.Desc(nx, ny, nz, 0.1, 0.05, 1e-3, 3, 4, flow, scalars=s, vectors=v)


def test_handles_initial_values_and_views():
    d = small_desc()
    with sf.Context(d) as ctx:
        assert list(ctx.scalar_handles) == [1, 2] and list(ctx.vector_handles) == [1]
        # initial state (reference :583-606)
        assert np.all(ctx.read(abi.VIEW_SCALAR_DATA, 1) == 0.5)
        assert np.all(ctx.read(abi.VIEW_SCALAR_DATA, 2) == 1.5)
        assert np.all(ctx.read(abi.VIEW_SCALAR_SOURCE, 1) == 0.0)
        fx, fy, fz = ctx.read(abi.VIEW_VECTOR, 1)
        assert np.all(fx == np.float32(0.1)) and np.all(fy == np.float32(0.2)) and np.all(fz == np.float32(0.3))
        for kind in (abi.VIEW_PRESSURE, abi.VIEW_DIVERGENCE, abi.VIEW_VELOCITY_MAGNITUDE):
            assert np.all(ctx.read(kind) == 0.0)
        v = ctx.get_view(abi.VIEW_VELOCITY)
        assert v.layout == abi.LAYOUT_F32_3D_SOA3 and (v.nx, v.ny, v.nz) == (10, 8, 6)
        assert v.row_stride_bytes == 40 and v.slice_stride_bytes == 320
        assert len({v.data0, v.data1, v.data2}) == 3
        p = ctx.get_view(abi.VIEW_PRESSURE)
        assert p.layout == abi.LAYOUT_F32_3D and p.data1 is None
        # views are stable library-owned pointers across steps
        ctx.step(2)
        assert ctx.get_view(abi.VIEW_PRESSURE).data0 == p.data0


def test_bad_handles_and_kinds():
    d = small_desc()
    with sf.Context(d) as ctx:
        lib, h = ctx.lib, ctx.handle
        for bad in (0, 3):
            assert lib.stable_fluids_update_scalar_field_cuda(h, bad, None) == abi.RESULT_BACKEND_FAILURE
            assert lib.stable_fluids_update_scalar_field_source_cuda(h, bad, None) == abi.RESULT_BACKEND_FAILURE
        for bad in (0, 2):
            assert lib.stable_fluids_update_vector_field_cuda(h, bad, None, None, None) == abi.RESULT_BACKEND_FAILURE
        view = abi.View()
        for req in (abi.ViewRequest(abi.VIEW_SCALAR_DATA, 0, 0, None), abi.ViewRequest(abi.VIEW_SCALAR_SOURCE, 3, 0, None),
                    abi.ViewRequest(abi.VIEW_VECTOR, 0, 2, None), abi.ViewRequest(7, 0, 0, None)):
            assert lib.stable_fluids_get_view_cuda(h, C.byref(req), C.byref(view)) == abi.RESULT_BACKEND_FAILURE
        assert lib.stable_fluids_ext_run_stage_cuda(h, 99, 1) == abi.RESULT_BACKEND_FAILURE


def test_updates_copy_and_null_zero_fills():
    d = small_desc()
    rng = np.random.default_rng(3)
    with sf.Context(d) as ctx:
        a = rng.random(d.shape).astype(np.float32)
        ctx.update_scalar_field(2, ctx.upload(a))
        ctx.update_scalar_field_source(1, ctx.upload(a * 2))
        fx, fy, fz = (rng.random(d.shape).astype(np.float32) for _ in range(3))
        ctx.update_vector_field(1, ctx.upload(fx), None, ctx.upload(fz))
        assert np.array_equal(ctx.read(abi.VIEW_SCALAR_DATA, 2), a)
        assert np.array_equal(ctx.read(abi.VIEW_SCALAR_SOURCE, 1), a * 2)
        gx, gy, gz = ctx.read(abi.VIEW_VECTOR, 1)
        assert np.array_equal(gx, fx) and np.all(gy == 0) and np.array_equal(gz, fz)
        ctx.update_scalar_field(2, None)
        assert np.all(ctx.read(abi.VIEW_SCALAR_DATA, 2) == 0)
        vx, vy, vz = (rng.standard_normal(d.shape).astype(np.float32) for _ in range(3))
        ctx.set_flow_velocity(ctx.upload(vx), ctx.upload(vy), ctx.upload(vz))
        got = ctx.read(abi.VIEW_VELOCITY)
        assert all(np.array_equal(g, w) for g, w in zip(got, (vx, vy, vz)))
        p = rng.standard_normal(d.shape).astype(np.float32)
        ctx.set_flow_pressure(ctx.upload(p))
        assert np.array_equal(ctx.read(abi.VIEW_PRESSURE), p)


def test_caller_stream_and_consumer_handoff():
    d = small_desc()
    stream, consumer = sf.hip.Stream(), sf.hip.Stream()
    try:
        with sf.Context(d, stream=stream) as ctx:
            assert ctx.stream == stream.handle.value
            ctx.step(3)
            view = ctx.get_view(abi.VIEW_SCALAR_DATA, 1, 0, consumer.handle)
            consumer.synchronize()
            host = sf.hip.to_host(view.data0, d.cells)
            assert np.all(np.isfinite(host))
            # same stream: no event needed
            ctx.get_view(abi.VIEW_PRESSURE, 0, 0, stream.handle)
    finally:
        consumer.destroy()
        stream.destroy()


def test_norms_match_host_reduction():
    d = small_desc(16, 12, 10)
    rng = np.random.default_rng(9)
    with sf.Context(d) as ctx:
        ctx.set_flow_velocity(*(ctx.upload(rng.standard_normal(d.shape)) for _ in range(3)))
        ctx.update_vector_field(1, *(ctx.upload(rng.standard_normal(d.shape)) for _ in range(3)))
        ctx.step(2)
        n = ctx.norms()
        div = ctx.read(abi.VIEW_DIVERGENCE).astype(np.float64)
        v = [a.astype(np.float64) for a in ctx.read(abi.VIEW_VELOCITY)]
        p = ctx.read(abi.VIEW_PRESSURE).astype(np.float64)
        assert n["divergence_l2"] == pytest.approx(np.sqrt(np.sum(div * div)), rel=1e-12)
        assert n["divergence_max"] == np.max(np.abs(div))
        assert n["velocity_l2"] == pytest.approx(np.sqrt(sum(np.sum(c * c) for c in v)), rel=1e-12)
        assert n["pressure_l2"] == pytest.approx(np.sqrt(np.sum(p * p)), rel=1e-12)


def test_stage_timing_hooks():
    """Stage hooks and their algorithmic bytes. The smoother-pass stages time the pass the
    step runs: with even nx, the black-cell pass (read b and the black cells, write the black
    cells: 8 B/cell); with odd nx, and for the full-iterate stages, 12 B/cell."""
    for n, pass_bytes in (((32, 16, 8), 8), ((31, 16, 8), 12)):
        d = small_desc(*n)
        with sf.Context(d) as ctx:
            for stage, per_cell in ((0, pass_bytes), (1, pass_bytes), (2, 24), (3, 36), (4, 12), (5, 8), (8, 12), (9, 12)):
                assert ctx.stage_bytes(stage) == per_cell * d.cells, (n, stage)
                ctx.run_stage(stage, 3)
            ctx.synchronize()


@pytest.mark.parametrize("smoother", [abi.SMOOTHER_RBGS, abi.SMOOTHER_JACOBI])
def test_projection_stage_hooks_recompute_the_step(smoother):
    """The PRESSURE_SOLVE and PROJECTION hooks enqueue exactly the step's sub-sequence: run
    after a step (same diffused velocity, same divergence) they reproduce the step's
    pressure, projected velocity and divergence bit for bit."""
    d = small_desc(32, 24, 16)
    rng = np.random.default_rng(4)
    with sf.Context(d, options=sf.ext_options(smoother=smoother)) as ctx:
        ctx.set_flow_velocity(*(ctx.upload(rng.standard_normal(d.shape)) for _ in range(3)))
        ctx.step(2)
        ref = {k: ctx.read(k) for k in (abi.VIEW_PRESSURE, abi.VIEW_DIVERGENCE)}
        ref_v = ctx.read(abi.VIEW_VELOCITY)
        for stage in (abi.STAGE_PRESSURE_SOLVE, abi.STAGE_PROJECTION):
            assert ctx.stage_bytes(stage) == 0
            ctx.run_stage(stage, 2)
            ctx.synchronize()
            for k, v in ref.items():
                assert np.array_equal(ctx.read(k), v), (stage, k)
            for a, b in zip(ctx.read(abi.VIEW_VELOCITY), ref_v):
                assert np.array_equal(a, b), stage


def test_many_contexts_create_destroy():
    for i in range(5):
        d = small_desc(8 + i, 8, 8)
        ctx = sf.Context(d)
        ctx.step(1)
        assert ctx.destroy() == abi.RESULT_OK
