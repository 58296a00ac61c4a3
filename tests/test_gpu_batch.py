"""Batched simulations (stable_fluids_ext_batch_*): one launch per stage for many contexts.

Pins: every context of a batch ends bit-identical to the same context stepped alone (on
uneven grids, mixed boundary types, several scalars/forces, RBGS and Jacobi); the
smoother tiling chosen per context (tile shape, z-chunk length, autotuner) never changes
a bit; the autotuner times every shape once and honours its cache file; the 64-plume
workload of uneven sizes at D=40, P=80 steps batched and autotuned bit-identical to each
simulation alone and, at 32^3, to the FP32 oracle.
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import sf, sf_oracle
from paritylib import Inputs, gpu_fields, oracle_fields

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


def test_jacobi_long_chunks_match_a_batch_of_one():
    """A stepped context runs its Jacobi passes on one workgroup per CU, here 128-plane z
    chunks; a batch of one runs the same passes on the default 64-plane plan: the fields
    must agree bit for bit."""
    n = (512, 512, 256)
    desc = desc_for(n, "walls", d_iters=6, p_iters=10)
    opts = sf.ext_options(smoother=abi.SMOOTHER_JACOBI)
    rng = np.random.default_rng(12)
    vel = [rng.standard_normal(desc.shape, dtype=np.float32) for _ in range(3)]

    def run(batched):
        ctx = sf.Context(desc, options=opts)
        try:
            ctx.set_flow_velocity(*(ctx.upload(v) for v in vel))
            if batched:
                with sf.Batch([ctx]) as b:
                    assert b.info(0)["z_chunk"] == 64
                    b.step(2)
                    sf.hip.synchronize()
            else:
                ctx.step(2)
            return gpu_fields(ctx, desc)
        finally:
            ctx.destroy()

    alone, batch = run(False), run(True)
    for k in alone:
        assert np.array_equal(alone[k], batch[k]), k


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
    # fixed chunk lengths, odd ones and one-plane chunks included, and (RBGS) every
    # black-cell tile shape, stepped as set
    fixed = make(opts)
    with sf.Batch([ctx for _, ctx in fixed]) as b:
        for i, zc in enumerate([2, 3, 5, 1, 9, 20]):
            if smoother == abi.SMOOTHER_RBGS:
                b.set_tiling(i, i % 5, zc)
                assert b.info(i)["tile"] == i % 5
            else:
                b.set_z_chunk(i, zc)
            assert b.info(i)["z_chunk"] == min(zc, SIMS[i][0][2])
        b.step(4)
        got = fields(fixed)
    destroy(fixed)
    assert_same(got, ref, "fixed tilings vs default tiling")
    tuned = make(opts)
    cache = str(tmp_path / "tiles.txt")
    with sf.Batch([ctx for _, ctx in tuned]) as b:
        shapes = b.autotune(cache, repeats=3)
        assert shapes == len(SIMS)
        b.step(4)
        got = fields(tuned)
        plans = [b.info(i) for i in range(len(SIMS))]
    destroy(tuned)
    assert_same(got, ref, "tuned vs default tiling")
    lines = [l for l in open(cache).read().split("\n") if l.strip()]
    assert lines[0] == "stable-fluids-tiling 2" and len(lines) == len(SIMS) + 1
    print("tuned plans", [(p["tile"], p["z_chunk"], p["chunks"]) for p in plans], flush=True)
    # a second batch takes every shape from the cache
    again = make(opts)
    with sf.Batch([ctx for _, ctx in again]) as b:
        assert b.autotune(cache, repeats=3) == 0
        got_plans = [b.info(i) for i in range(len(SIMS))]
        assert [(p["tile"], p["z_chunk"]) for p in got_plans] == [(p["tile"], p["z_chunk"]) for p in plans]
    destroy(again)


def digests(ctx, desc):
    return {k: hashlib.sha1(np.ascontiguousarray(v).tobytes()).hexdigest() for k, v in gpu_fields(ctx, desc).items()}


def test_64_uneven_plumes_batched_autotuned(tmp_path):
    """64 plumes with sizes drawn from {32^3, 48^3, 64^3, 96^3, 128^3, 128x64x64}, radius
    U[2, 8], force U[1, 10], D=40, P=80: batched and autotuned, 3 steps."""
    scs = sf.scenes.batch_plume_scenes(64, seed=0, diffuse_iterations=40, pressure_iterations=80)
    assert len({sc.desc.shape for sc in scs}) == 6
    alone = []
    for sc in scs:
        ctx = sf.Context(sc.desc)
        try:
            sf.scenes.seed_context(ctx, sc)
            ctx.step(3)
            alone.append(digests(ctx, sc.desc))
        finally:
            ctx.destroy()
    ctxs = []
    try:
        for sc in scs:
            ctxs.append(sf.Context(sc.desc))
            sf.scenes.seed_context(ctxs[-1], sc)
        with sf.Batch(ctxs) as b:
            tuned = b.autotune(str(tmp_path / "tiles.txt"), repeats=3)
            plans = {sc.desc.shape: (b.info(i)["tile"], b.info(i)["z_chunk"]) for i, sc in enumerate(scs)}
            b.step(3)
            sf.hip.synchronize()
            small = [(i, gpu_fields(ctxs[i], sc.desc)) for i, sc in enumerate(scs) if sc.desc.shape == (32, 32, 32)]
            batched = [digests(c, sc.desc) for c, sc in zip(ctxs, scs)]
            launches = b.info(0)["launches_per_step"]
    finally:
        for c in ctxs:
            c.destroy()
    print(f"tuned {tuned} shapes: {plans}; {launches} launches per batch step", flush=True)
    assert tuned == 6
    for i, (g, a) in enumerate(zip(batched, alone)):
        bad = [k for k in a if g[k] != a[k]]
        assert not bad, f"simulation {i} {scs[i].desc.shape}: batched differs from alone in {bad}"
    assert small, "no 32^3 simulation drawn"
    for i, got in small:
        sc = scs[i]
        sim = sf_oracle.OracleSim(abi, sc.desc, precision=32)
        try:
            sf.scenes.seed_oracle(sim, sc)
            sim.step(3)
            ref = oracle_fields(sim, sc.desc)
        finally:
            sim.destroy()
        bad = [k for k in ref if not np.array_equal(got[k], ref[k].astype(np.float64))]
        assert not bad, f"simulation {i} (32^3) differs from the FP32 oracle in {bad}"


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
