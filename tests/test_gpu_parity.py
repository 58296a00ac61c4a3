"""GPU step parity against the host-loop oracle, through the C ABI.

Two bars:
  * bit-exact against the FP32 restatement (same operations, same rounding): the HIP
    kernels compute every value with the reference's expression order and no FMA
    contraction, so every field must match to the last bit, at any step count;
  * norm-wise relative error max|gpu - o64| / max|o64| <= 1e-5 per field against the
    FP64 oracle, on horizons where FP32 rounding has not been amplified. The 64^3 plume
    at dt=0.1, force 5 is chaotic (the FP32 and FP64 oracles themselves drift apart
    ~10x every 2-3 steps, 40% by step 20), so its FP64 bar is checked for 3 steps and
    for 100 steps on the same scene with force 0.5, which is well conditioned.
"""
import zlib

import numpy as np
import pytest

from conftest import sf, sf_oracle
from paritylib import Inputs, gpu_fields, normwise, report, run_gpu, run_oracle

abi = sf.abi
scenes = sf.scenes
pytestmark = pytest.mark.gpu
TOL64 = 1e-5


def plume_seeder(scene):
    return (lambda ctx: scenes.seed_context(ctx, scene)), (lambda sim: scenes.seed_oracle(sim, scene))


def assert_bitexact(gpu, ref, tag):
    errs = normwise(gpu, ref)
    report(tag, errs)
    bad = {k: v for k, v in errs.items() if v != 0.0}
    assert not bad, f"{tag}: not bit-exact: {bad}"


def assert_fp64(gpu, ref, tag, tol=TOL64):
    errs = normwise(gpu, ref)
    report(tag, errs)
    bad = {k: v for k, v in errs.items() if v > tol}
    assert not bad, f"{tag}: above {tol}: {bad}"


def test_plume64_100_steps_bitexact_fp32_oracle():
    """The 64^3 plume (dt 0.1, nu=kappa=1e-4, D=20, P=40, free-slip walls), 100 steps."""
    scene = scenes.plume_scene(64)
    g_seed, o_seed = plume_seeder(scene)
    gpu = run_gpu(scene.desc, g_seed, 100)
    ref32 = run_oracle(scene.desc, o_seed, 100, 32)
    assert_bitexact(gpu, ref32, "plume64 100 steps vs fp32 oracle")
    assert np.max(gpu["scalar1"]) > 1.0  # the plume is there


def test_plume64_3_steps_fp64():
    scene = scenes.plume_scene(64)
    g_seed, o_seed = plume_seeder(scene)
    assert_fp64(run_gpu(scene.desc, g_seed, 3), run_oracle(scene.desc, o_seed, 3, 64), "plume64 3 steps vs fp64")


def test_plume64_gentle_100_steps_fp64():
    """Same scene and solver settings with force 0.5: FP32 stays within 1e-5 of FP64."""
    scene = scenes.plume_scene(64)
    src, fx, fy, fz = scenes.plume_fields(scene.desc.shape, (32, 6, 32), 4, 100.0, 0.5)
    scene.force = (fx, fy, fz)
    g_seed, o_seed = plume_seeder(scene)
    assert_fp64(run_gpu(scene.desc, g_seed, 100), run_oracle(scene.desc, o_seed, 100, 64), "plume64 force0.5 100 steps vs fp64")


def mixed_desc(n=(12, 10, 8), d_iters=6, p_iters=8, viscosity=2e-3, diffusion=1e-3, dissipation=0.0):
    nx, ny, nz = n
    wall_v = [(0, 0, 0), (0, 0, 0), (0.3, 0.0, 0.1), (0, 0, 0), (0.05, -0.2, 0.0), (0.1, 0.0, 0.0)]
    flow = abi.flow_boundary([abi.FLOW_PERIODIC, abi.FLOW_PERIODIC, abi.FLOW_NO_SLIP, abi.FLOW_OUTFLOW,
                              abi.FLOW_FREE_SLIP, abi.FLOW_NO_SLIP], velocity=wall_v, pressure=[0, 0, 0, 0.5, 0, 0])
    s1 = abi.scalar_boundary([abi.SCALAR_PERIODIC, abi.SCALAR_PERIODIC, abi.SCALAR_FIXED, abi.SCALAR_FIXED,
                              abi.SCALAR_ZERO_FLUX, abi.SCALAR_FIXED], values=[0, 0, 1.0, 0.0, 0, -0.5])
    s2 = abi.scalar_boundary([abi.SCALAR_FIXED, abi.SCALAR_ZERO_FLUX, abi.SCALAR_PERIODIC, abi.SCALAR_PERIODIC,
                              abi.SCALAR_ZERO_FLUX, abi.SCALAR_ZERO_FLUX], values=[2.0, 0, 0, 0, 0, 0])
    scalars = [abi.scalar_desc(diffusion, dissipation, 0.25, s1), abi.scalar_desc(diffusion * 3, 0.7, -1.0, s2)]
    vectors = [abi.force_desc(0.1, 0.0, -0.2), abi.force_desc(0.0, 0.3, 0.0)]
    return abi.Desc(nx, ny, nz, 1.0 / 16, 0.05, viscosity, d_iters, p_iters, flow, scalars=scalars, vectors=vectors)


EDGE_CASES = {
    "mixed_bc": dict(),
    "odd_sizes_walls": dict(n=(7, 9, 5)),
    "odd_periodic_x": dict(n=(7, 6, 5)),
    "slab_nz1": dict(n=(16, 12, 1)),
    "line_nx1": dict(n=(1, 9, 6)),
    "wide_nx": dict(n=(130, 3, 4)),
    "nu0_kappa0_quirk": dict(viscosity=0.0, diffusion=0.0),
    "dissipation": dict(dissipation=0.35),
    "no_iterations": dict(d_iters=0, p_iters=0),
    "odd_iterations": dict(d_iters=5, p_iters=7),
}


@pytest.mark.parametrize("case", sorted(EDGE_CASES))
def test_edge_cases_bitexact_and_fp64(case):
    desc = mixed_desc(**EDGE_CASES[case])
    inputs = Inputs(desc, seed=zlib.crc32(case.encode()) % 1000, velocity_scale=0.5)
    gpu = run_gpu(desc, inputs.seed_gpu, 6)
    assert_bitexact(gpu, run_oracle(desc, inputs.seed_oracle, 6, 32), f"{case} 6 steps vs fp32")
    # FP64 bar on box-filtered inputs: with white noise per cell, FP32 world positions
    # (resolution ~8e-6 cells at x = 8) times O(1)-per-cell gradients already move the
    # FP32 oracle itself 1.6e-5 away from FP64 in one step
    smooth = Inputs(desc, seed=zlib.crc32(case.encode()) % 1000, velocity_scale=0.5, smooth=True)
    gpu1 = run_gpu(desc, smooth.seed_gpu, 1)
    assert_fp64(gpu1, run_oracle(desc, smooth.seed_oracle, 1, 64), f"{case} smooth inputs 1 step vs fp64")


def test_all_periodic_large_cfl_bitexact():
    """Backtraces that wrap several times around periodic axes (fmodf path)."""
    flow = abi.flow_boundary([abi.FLOW_PERIODIC] * 6)
    sb = abi.scalar_boundary([abi.SCALAR_PERIODIC] * 6)
    desc = abi.Desc(16, 8, 12, 1.0 / 16, 0.5, 1e-3, 4, 6, flow, scalars=[abi.scalar_desc(1e-3, 0.0, 0.0, sb)], vectors=[abi.force_desc()])
    inputs = Inputs(desc, seed=5, velocity_scale=6.0)
    gpu = run_gpu(desc, inputs.seed_gpu, 4)
    assert_bitexact(gpu, run_oracle(desc, inputs.seed_oracle, 4, 32), "periodic large cfl vs fp32")


def test_jacobi_smoother_bitexact():
    desc = mixed_desc(d_iters=7, p_iters=9)
    inputs = Inputs(desc, seed=11, velocity_scale=0.5)
    opts = sf.ext_options(smoother=abi.SMOOTHER_JACOBI)
    gpu = run_gpu(desc, inputs.seed_gpu, 5, options=opts)
    assert_bitexact(gpu, run_oracle(desc, inputs.seed_oracle, 5, 32, smoother=sf_oracle.JACOBI), "jacobi vs fp32")
    smooth = Inputs(desc, seed=11, velocity_scale=0.5, smooth=True)
    gpu1 = run_gpu(desc, smooth.seed_gpu, 1, options=opts)
    assert_fp64(gpu1, run_oracle(desc, smooth.seed_oracle, 1, 64, smoother=sf_oracle.JACOBI), "jacobi smooth inputs 1 step vs fp64")


def walls_desc(n=(12, 10, 8), d_iters=7, p_iters=9):
    """No-slip x- wall with a moving lid (reflect ghosts), outflow y+ and z+ (constant
    pressure ghosts), free-slip elsewhere; fixed-value and zero-flux scalar faces."""
    nx, ny, nz = n
    wall_v = [(0.0, 0.2, -0.1), (0, 0, 0), (0, 0, 0), (0, 0, 0), (0.05, -0.2, 0.0), (0, 0, 0)]
    flow = abi.flow_boundary([abi.FLOW_NO_SLIP, abi.FLOW_FREE_SLIP, abi.FLOW_FREE_SLIP, abi.FLOW_OUTFLOW,
                              abi.FLOW_NO_SLIP, abi.FLOW_OUTFLOW], velocity=wall_v, pressure=[0, 0, 0, 0.5, 0, -0.25])
    s1 = abi.scalar_boundary([abi.SCALAR_FIXED, abi.SCALAR_ZERO_FLUX, abi.SCALAR_ZERO_FLUX, abi.SCALAR_FIXED,
                              abi.SCALAR_FIXED, abi.SCALAR_ZERO_FLUX], values=[1.0, 0, 0, -0.5, 0.25, 0])
    scalars = [abi.scalar_desc(1e-3, 0.2, 0.25, s1)]
    vectors = [abi.force_desc(0.1, 0.0, -0.2)]
    return abi.Desc(nx, ny, nz, 1.0 / 16, 0.05, 2e-3, d_iters, p_iters, flow, scalars=scalars, vectors=vectors)


def periodic_yz_desc(n=(12, 10, 8), d_iters=7, p_iters=9):
    """Periodic y and z, a no-slip x- wall with a moving lid and an outflow x+ face."""
    nx, ny, nz = n
    wall_v = [(0.0, 0.15, -0.1), (0, 0, 0), (0, 0, 0), (0, 0, 0), (0, 0, 0), (0, 0, 0)]
    flow = abi.flow_boundary([abi.FLOW_NO_SLIP, abi.FLOW_OUTFLOW, abi.FLOW_PERIODIC, abi.FLOW_PERIODIC,
                              abi.FLOW_PERIODIC, abi.FLOW_PERIODIC], velocity=wall_v, pressure=[0, 0.3, 0, 0, 0, 0])
    s1 = abi.scalar_boundary([abi.SCALAR_FIXED, abi.SCALAR_ZERO_FLUX, abi.SCALAR_PERIODIC, abi.SCALAR_PERIODIC,
                              abi.SCALAR_PERIODIC, abi.SCALAR_PERIODIC], values=[0.75, 0, 0, 0, 0, 0])
    scalars = [abi.scalar_desc(1e-3, 0.0, 0.25, s1)]
    vectors = [abi.force_desc(0.1, -0.05, 0.2)]
    return abi.Desc(nx, ny, nz, 1.0 / 16, 0.05, 2e-3, d_iters, p_iters, flow, scalars=scalars, vectors=vectors)


# smoother tile shapes: 128 x 16 cells per tile, z chunks of the default plan
TILE_CASES = {
    "periodic_one_full_tile": (mixed_desc, (128, 6, 3)),     # halo columns wrap
    "periodic_partial_tile": (mixed_desc, (130, 3, 4)),      # pair at nx carries the wrapped cell
    "walls_two_full_tiles": (walls_desc, (256, 18, 5)),      # wall halo columns, partial y tile
    "walls_partial_tile": (walls_desc, (130, 3, 4)),         # pair at nx carries the x+ ghost
    "walls_odd_nx": (walls_desc, (9, 7, 6)),
    "walls_slab_nz1": (walls_desc, (16, 12, 1)),             # one-plane chunks
    # black-cell passes: a 2-pair x tile and a 2-row y tile past full ones, 10 z chunks
    "walls_multi_tile_chunks": (walls_desc, (132, 34, 40)),
    "periodic_yz_partial": (periodic_yz_desc, (136, 18, 20)),  # wrapped halo rows and planes
    # periodic y and z across many z chunks of the z-marching projection and divergence
    # (4-plane chunks at this size, the last one 2 planes; their z windows wrap at the
    # first and last chunk), partial x and y tiles
    "periodic_yz_z_chunks": (periodic_yz_desc, (70, 20, 70)),
    "mixed_two_cells": (mixed_desc, (4, 2, 2)),                # every cell on a face
    "walls_two_cells": (walls_desc, (4, 2, 2)),
    # 32 tiles of 128 x 16: Jacobi passes run 8-plane chunks (one workgroup per CU),
    # RBGS ones their own plan
    "walls_jacobi_long_chunks": (walls_desc, (512, 128, 64)),
}


@pytest.mark.parametrize("smoother", ["rbgs", "jacobi"])
@pytest.mark.parametrize("case", sorted(TILE_CASES))
def test_tile_shapes_bitexact(case, smoother):
    make, n = TILE_CASES[case]
    desc = make(n=n)
    inputs = Inputs(desc, seed=zlib.crc32(case.encode()) % 1000, velocity_scale=0.5)
    kind = abi.SMOOTHER_JACOBI if smoother == "jacobi" else abi.SMOOTHER_RBGS
    ref_kind = sf_oracle.JACOBI if smoother == "jacobi" else sf_oracle.RBGS
    gpu = run_gpu(desc, inputs.seed_gpu, 3, options=sf.ext_options(smoother=kind))
    assert_bitexact(gpu, run_oracle(desc, inputs.seed_oracle, 3, 32, smoother=ref_kind), f"{case} {smoother} 3 steps vs fp32")


# (case, tile shape, z chunk override): each black-cell tile width and height, a chunk
# override that splits nz unevenly, and tiles past the grid on both axes
TILING_CASES = [
    ("walls_multi_tile_chunks", abi.TILE_128X16, 0),
    ("walls_multi_tile_chunks", abi.TILE_64X32, 6),
    ("walls_multi_tile_chunks", abi.TILE_64X16, 4),
    ("walls_multi_tile_chunks", abi.TILE_32X32, 10),
    ("walls_multi_tile_chunks", abi.TILE_32X16, 0),
    ("periodic_yz_partial", abi.TILE_64X32, 0),
    ("periodic_yz_partial", abi.TILE_32X16, 6),
    ("mixed_two_cells", abi.TILE_32X16, 0),
    ("periodic_partial_tile", abi.TILE_32X32, 4),
]


@pytest.mark.parametrize("case,tile,zc", TILING_CASES)
def test_smoother_tilings_bitexact(case, tile, zc):
    """Every black-cell tile shape steps bit-exactly against the FP32 oracle."""
    make, n = TILE_CASES[case]
    desc = make(n=n)
    inputs = Inputs(desc, seed=zlib.crc32(case.encode()) % 1000, velocity_scale=0.5)
    ctx = sf.Context(desc)
    try:
        ctx.set_smoother_tiling(tile, zc)
        got_tile, got_zc = ctx.smoother_tiling()
        assert got_tile == tile and (got_zc == min(zc, n[2]) if zc else got_zc > 0)
        inputs.seed_gpu(ctx)
        ctx.step(3)
        gpu = gpu_fields(ctx, desc)
    finally:
        ctx.destroy()
    assert_bitexact(gpu, run_oracle(desc, inputs.seed_oracle, 3, 32), f"{case} tile {tile} zc {zc} 3 steps vs fp32")


def test_single_context_autotune_bitexact(tmp_path):
    """The tiling autotuner picks a plan for one context, caches it, and the tuned context
    steps bit-exactly against the FP32 oracle; a second context takes the cached plan."""
    make, n = TILE_CASES["walls_multi_tile_chunks"]
    desc = make(n=n)
    inputs = Inputs(desc, seed=7, velocity_scale=0.5)
    cache = str(tmp_path / "tiles.txt")
    ctx = sf.Context(desc)
    try:
        plan = ctx.autotune(cache, repeats=3)
        assert ctx.smoother_tiling() == plan
        inputs.seed_gpu(ctx)
        ctx.step(3)
        gpu = gpu_fields(ctx, desc)
    finally:
        ctx.destroy()
    print("tuned tiling", plan, flush=True)
    lines = [l for l in open(cache).read().split("\n") if l.strip()]
    assert lines[0] == "stable-fluids-tiling 2" and lines[1].split()[:5] == [str(v) for v in (*n, abi.SMOOTHER_RBGS, 0)]
    assert_bitexact(gpu, run_oracle(desc, inputs.seed_oracle, 3, 32), "autotuned 3 steps vs fp32")
    again = sf.Context(desc)
    try:
        assert again.autotune(cache, repeats=3) == plan
    finally:
        again.destroy()


def test_plume256_noise_one_step_bitexact():
    """The 256^3 throughput workload (seeded divergent initial velocity, D=40, P=80), one step."""
    scene = scenes.plume_scene(256, diffuse_iterations=40, pressure_iterations=80, noise_seed=0)
    g_seed, o_seed = plume_seeder(scene)
    gpu = run_gpu(scene.desc, g_seed, 1)
    assert_bitexact(gpu, run_oracle(scene.desc, o_seed, 1, 32), "plume256 1 step vs fp32")


def test_determinism_two_runs_bitwise():
    scene = scenes.plume_scene(64)
    g_seed, _ = plume_seeder(scene)
    a = run_gpu(scene.desc, g_seed, 20)
    b = run_gpu(scene.desc, g_seed, 20)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_reference_benchmark_scene_bitexact():
    """The reference benchmark's scene (benchmark.cu:72-155): periodic x/z, no-slip floor,
    outflow top, dissipation 0.35, swirling ellipsoidal plume, h = 0.01 (not a power of
    two, so no trilinear weight is exact), at 32x40x32."""
    scene = scenes.reference_benchmark_scene(32, 40, 32, diffuse_iterations=6, pressure_iterations=12)
    g_seed, o_seed = plume_seeder(scene)
    assert_bitexact(run_gpu(scene.desc, g_seed, 12), run_oracle(scene.desc, o_seed, 12, 32), "benchmark scene 12 steps vs fp32")
    assert_fp64(run_gpu(scene.desc, g_seed, 2), run_oracle(scene.desc, o_seed, 2, 64), "benchmark scene 2 steps vs fp64")


def test_checkpoint_resume_bitwise():
    """Saving velocity and density after 10 steps and restoring them into a fresh context
    continues the run bit for bit (pressure restarts from zero every step)."""
    scene = scenes.plume_scene(48, diffuse_iterations=8, pressure_iterations=16)
    g_seed, _ = plume_seeder(scene)
    straight = run_gpu(scene.desc, g_seed, 20)
    with sf.Context(scene.desc) as a:
        g_seed(a)
        a.step(10)
        vel = a.read(abi.VIEW_FLOW_VELOCITY)
        rho = a.read(abi.VIEW_SCALAR_FIELD_DATA, 1)
    with sf.Context(scene.desc) as b:
        g_seed(b)
        b.set_flow_velocity(*(b.upload(c) for c in vel))
        b.update_scalar_field(1, b.upload(rho))
        b.step(10)
        resumed = {"scalar1": b.read(abi.VIEW_SCALAR_FIELD_DATA, 1).astype(np.float64)}
        for name, arr in zip(("vx", "vy", "vz"), b.read(abi.VIEW_FLOW_VELOCITY)):
            resumed[name] = arr.astype(np.float64)
    for k, v in resumed.items():
        assert np.array_equal(v, straight[k]), k


def test_reference_benchmark_driver_output():
    """stable-fluids-benchmark keeps the reference driver's flags and output lines."""
    import os
    import re
    import subprocess
    exe = os.path.join(os.path.dirname(sf.runtime.LIB_PATH), "stable-fluids-benchmark")
    out = subprocess.run([exe, "--nx", "32", "--ny", "40", "--nz", "32", "--warmup", "2", "--steps", "4"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().splitlines()
    assert lines[0].startswith("benchmark grid=32x40x32 warmup=2 steps=4 ")
    m = re.match(r"timing total_ms=([\d.]+) avg_step_ms=([\d.]+) mlups=([\d.]+) total_density=([-\d.]+) peak_density=([-\d.]+)", lines[1])
    assert m and float(m.group(3)) > 0 and float(m.group(5)) > 0
