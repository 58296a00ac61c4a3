"""The allocation-free norms entry point (stable_fluids_ext_field_norms_noalloc_cuda)
against the context entry point it mirrors (stable_fluids_ext_field_norms_cuda): two calls
in a row through the C ABI, the second on a smaller grid, each bit-identical."""
import ctypes as C

import numpy as np
import pytest

from conftest import sf

abi = sf.abi
pytestmark = pytest.mark.gpu


def _stepped_context(n, steps):
    """A plume context with a seeded smoothed-noise velocity, stepped `steps` times, so that
    divergence, velocity and pressure are all non-zero."""
    scene = sf.scenes.plume_scene(n, diffuse_iterations=6, pressure_iterations=12, noise_seed=n)
    ctx = sf.Context(scene.desc)
    sf.scenes.seed_context(ctx, scene)
    ctx.step(steps)
    return ctx


def _noalloc(ctx, out_dev, stream):
    div = ctx.get_view(abi.VIEW_DIVERGENCE).data0
    v = ctx.get_view(abi.VIEW_VELOCITY)
    p = ctx.get_view(abi.VIEW_PRESSURE).data0
    r = ctx.lib.stable_fluids_ext_field_norms_noalloc_cuda(div, v.data0, v.data1, v.data2, p, ctx.cells, stream, out_dev.value)
    assert r == abi.RESULT_OK
    host = sf.runtime.ExtNorms()
    sf.hip.synchronize()
    sf.hip.check(sf.hip.lib().hipMemcpy(C.byref(host), C.c_void_p(out_dev.value), C.c_size_t(C.sizeof(host)), 2), "hipMemcpy D2H")
    return {k: getattr(host, k) for k, _ in host._fields_}


@pytest.mark.parametrize("sizes", [(72, 20)])
def test_noalloc_norms_match_context_norms_twice(sizes):
    out_dev = sf.hip.DeviceBuffer(C.sizeof(sf.runtime.ExtNorms))
    for n in sizes:  # the second grid is smaller: fewer partial sums than the first call left
        print(f"no-alloc norms: plume {n}^3", flush=True)
        ctx = _stepped_context(n, 2)
        try:
            ref = ctx.norms()
            got = _noalloc(ctx, out_dev, ctx.lib.stable_fluids_ext_get_stream_cuda(ctx.handle))
            assert all(v > 0.0 for v in ref.values()), ref
            for k, v in ref.items():
                assert got[k] == v, (n, k, got[k], v)
        finally:
            ctx.destroy()


def test_noalloc_norms_reject_bad_arguments():
    lib = sf.load_library()
    out_dev = sf.hip.DeviceBuffer(C.sizeof(sf.runtime.ExtNorms))
    f = sf.hip.DeviceBuffer(64 * 4)
    args = [f.value] * 5
    assert lib.stable_fluids_ext_field_norms_noalloc_cuda(*args, 0, None, out_dev.value) == abi.RESULT_BACKEND_FAILURE
    assert lib.stable_fluids_ext_field_norms_noalloc_cuda(*args, 64, None, None) == abi.RESULT_BACKEND_FAILURE
    assert lib.stable_fluids_ext_field_norms_noalloc_cuda(None, *args[1:], 64, None, out_dev.value) == abi.RESULT_BACKEND_FAILURE
    # 64 cells of zeros on the null stream
    sf.hip.to_device(f.value, np.zeros(64, np.float32))
    assert lib.stable_fluids_ext_field_norms_noalloc_cuda(*args, 64, None, out_dev.value) == abi.RESULT_OK
    sf.hip.synchronize()
    assert np.all(sf.hip.to_host(out_dev.value, 4, np.float64) == 0.0)
