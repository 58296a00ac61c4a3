"""Batched simulations (stable_fluids_ext_batch_*): one launch per stage for many contexts.

Pins: every context of a batch ends bit-identical to the same context stepped alone (on
uneven grids, mixed boundary types, several scalars/forces, RBGS and Jacobi); the
smoother tiling chosen per context (z-chunk length, autotuner) never changes a bit; the
autotuner times every shape once and honours its cache file.
"""
import os

import numpy as np
import pytest

from conftest import sf
from paritylib import Inputs, gpu_fields

abi = sf.abi
pytestmark = pytest.mark.gpu


def desc_for(n, kind, d_iters=6, p_iters=10):
    nx, ny, nz = n
    if kind == "walls":
        flow = abi.flow_boundary([abi.FLOW_FREE_SLIP] * 6)
        sb = abi.scalar_boundary([abi.SCALAR_ZERO_FLUX] * 6)
    elif kind == "mixed":
        flow = abi.flow_boundary([abi.FLOW_PERIODIC, abi.FLOW_PERIODIC, abi.FLOW_NO_SLIP, abi.FLOW_OUTFLOW,
                                  abi.FLOW_FREE_SLIP, abi.FLOW_NO_SLIP], velocity=[(0, 0, 0)] * 2 + [(0.3, 0, 0.1)] + [(0, 0, 0)] * 3,
                                 pressure=[0, 0, 0, 0.5, 0, 0])
        sb = abi.scalar_boundary([abi.SCALAR_PERIODIC, abi.SCALAR_PERIODIC, abi.SCALAR_FIXED, abi.SCALAR_FIXED,
                                  abi.SCALAR_ZERO_FLUX, abi.SCALAR_FIXED], values=[0, 0, 1.0, 0, 0, -0.5])
    else:  # periodic box
        flow = abi.flow_boundary([abi.FLOW_PERIODIC] * 6)
        sb = abi.scalar_boundary([abi.SCALAR_PERIODIC] * 6)
    scalars = [abi.scalar_desc(2e-3, 0.3, 0.1, sb), abi.scalar_desc(1e-3, 0.0, 0.0, sb)]
    vectors = [abi.force_desc(0.0, 0.5, 0.0), abi.force_desc(0.1, 0.0, 0.0)]
    return abi.Desc(nx, ny, nz, 1.0 / max(n), 0.05, 1e-3, d_iters, p_iters, flow, scalars=scalars, vectors=vectors)


SIMS = [((32, 32, 32), "walls"), ((48, 40, 24), "mixed"), ((64, 16, 12), "periodic"), ((24, 24, 24), "mixed"),
        ((17, 33, 9), "walls"), ((130, 8, 20), "mixed")]


def make(options=None):
    out = []
    for i, (n, kind) in enumerate(SIMS):
        desc = desc_for(n, kind)
        inputs = Inputs(desc, seed=100 + i, velocity_scale=0.5, smooth=True)
        ctx = sf.Context(desc, options=options)
        inputs.seed_gpu(ctx)
        out.append((desc, ctx))
    return out


def fields(sims):
    sf.hip.synchronize()
    return [gpu_fields(ctx, desc) for desc, ctx in sims]


def destroy(sims):
    for _, ctx in sims:
        ctx.destroy()


def assert_same(a, b, tag):
    for i, (fa, fb) in enumerate(zip(a, b)):
        for k in fb:
            assert np.array_equal(fa[k], fb[k]), f"{tag}: sim {i} field {k}"


@pytest.mark.parametrize("smoother", [abi.SMOOTHER_RBGS, abi.SMOOTHER_JACOBI])
def test_batch_equals_each_context_alone(smoother):
    opts = sf.ext_options(smoother=smoother)
    alone = make(opts)
    for _, ctx in alone:
        ctx.step(5)
    ref = fields(alone)
    destroy(alone)
    together = make(opts)
    with sf.Batch([ctx for _, ctx in together]) as b:
        info = b.info(0)
        b.step(5)
        got = fields(together)
    destroy(together)
    print(f"smoother {smoother}: {info['launches_per_step']} launches per batch step for {len(SIMS)} simulations", flush=True)
    assert_same(got, ref, "batch vs alone")


@pytest.mark.parametrize("smoother", [abi.SMOOTHER_RBGS, abi.SMOOTHER_JACOBI])
def test_tiling_never_changes_results(tmp_path, smoother):
    opts = sf.ext_options(smoother=smoother)
    base = make(opts)
    with sf.Batch([ctx for _, ctx in base]) as b:
        b.step(4)
        ref = fields(base)
    destroy(base)
    tuned = make(opts)
    cache = str(tmp_path / "tiles.txt")
    with sf.Batch([ctx for _, ctx in tuned]) as b:
        for i, zc in enumerate([2, 3, 5, 1, 9, 20]):
            b.set_z_chunk(i, zc)
            assert b.info(i)["z_chunk"] == min(zc, SIMS[i][0][2])
        shapes = b.autotune(cache, repeats=3)
        assert shapes == len(SIMS)
        b.step(4)
        got = fields(tuned)
        plans = [b.info(i) for i in range(len(SIMS))]
    destroy(tuned)
    assert_same(got, ref, "tuned vs default tiling")
    lines = open(cache).read().split("\n")
    assert len([l for l in lines if l.strip()]) == len(SIMS)
    print("tuned plans", [(p["z_chunk"], p["chunks"]) for p in plans], flush=True)
    # a second batch takes every shape from the cache
    again = make(opts)
    with sf.Batch([ctx for _, ctx in again]) as b:
        assert b.autotune(cache, repeats=3) == 0
        assert [b.info(i)["z_chunk"] for i in range(len(SIMS))] == [p["z_chunk"] for p in plans]
    destroy(again)


def test_incompatible_contexts_are_rejected():
    a = sf.Context(desc_for((16, 16, 16), "walls"))
    b = sf.Context(desc_for((16, 16, 16), "walls", p_iters=12))
    c = sf.Context(desc_for((16, 16, 16), "walls"), options=sf.ext_options(iterate_storage=abi.STORAGE_F16))
    try:
        for group in ([a, b], [a, c], [a, a]):
            with pytest.raises(sf.StableFluidsError):
                sf.Batch(group)
    finally:
        for x in (a, b, c):
            x.destroy()
