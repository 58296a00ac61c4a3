"""Two RBGS passes per launch (STABLE_FLUIDS_EXT_FLAG_PASS_PAIRS) against the oracle.

The pair kernel relaxes five half-sweeps (old black cells, red and black of iteration 1,
red and black of iteration 2) in one z march with their own LDS halos and ghost
positions, so the shapes here cover what differs from the single pass: chunk halos four
planes deep (many z chunks), ghost rows and x ghost pairs one to four cells from the tile,
partial tiles, a one-plane slab, periodic x and z, odd pass counts (one single pass, then
pairs) and even ones (the first pair starts from zeros), and the learned warm start,
whose skipped launches now come in pairs.
"""
import zlib

import numpy as np
import pytest

from conftest import sf, sf_oracle
from oracle import unet_oracle as uo
from paritylib import Inputs, gpu_fields, run_gpu, run_oracle
from test_gpu_parity import assert_bitexact, mixed_desc, walls_desc

abi = sf.abi
scenes = sf.scenes
pytestmark = pytest.mark.gpu

PAIRS = sf.ext_options(pass_pairs=True)

CASES = {
    "walls_two_full_tiles_odd_passes": (walls_desc, dict(n=(256, 18, 5))),
    "walls_partial_tile": (walls_desc, dict(n=(130, 3, 4))),
    "walls_slab_nz1": (walls_desc, dict(n=(16, 12, 1))),
    "walls_many_chunks_even_passes": (walls_desc, dict(n=(64, 20, 40), d_iters=6, p_iters=8)),
    "periodic_x_ghost_rows": (mixed_desc, dict(n=(128, 21, 14))),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_pass_pairs_bitexact(case):
    make, kw = CASES[case]
    desc = make(**kw)
    inputs = Inputs(desc, seed=zlib.crc32(case.encode()) % 1000, velocity_scale=0.5)
    gpu = run_gpu(desc, inputs.seed_gpu, 3, options=PAIRS)
    assert_bitexact(gpu, run_oracle(desc, inputs.seed_oracle, 3, 32, smoother=sf_oracle.RBGS), f"{case} pass pairs 3 steps vs fp32")


def test_pass_pairs_periodic_z_benchmark_scene():
    """Periodic x and z (nz >= 12 takes the pair kernel), no-slip floor, outflow top."""
    scene = scenes.reference_benchmark_scene(32, 40, 32, diffuse_iterations=6, pressure_iterations=12)
    g_seed = lambda ctx: scenes.seed_context(ctx, scene)  # noqa: E731
    pairs = run_gpu(scene.desc, g_seed, 6, options=PAIRS)
    single = run_gpu(scene.desc, g_seed, 6)
    for k in single:
        assert np.array_equal(pairs[k], single[k]), k


def plume(p_iters=12):
    return scenes.plume_scene(32, diffuse_iterations=4, pressure_iterations=p_iters, noise_seed=1)


def run_learned(scene, params, k, pairs):
    opts = sf.ext_options(pressure_warm_start=abi.WARM_START_LEARNED, warm_start_iterations=k, pass_pairs=pairs)
    with sf.Context(scene.desc, options=opts) as ctx:
        ctx.set_pressure_model(params)
        scenes.seed_context(ctx, scene)
        ctx.step(3)
        stats = ctx.warm_start_stats()
        return gpu_fields(ctx, scene.desc), stats


@pytest.mark.parametrize("k", [0, 4, 5, 6])
@pytest.mark.parametrize("net", ["accepted", "fallback"])
def test_learned_warm_start_pairs_bitwise(k, net):
    """Accepted guesses with C = P - K skipped passes (C even and odd) and fallbacks."""
    scene = plume()
    if net == "accepted":
        layers = uo.init_params(4)
        w, b = layers[-1]
        layers[-1] = (w * 0.05, b)
        params = uo.join_params(layers).astype(np.float32)
    else:
        params = uo.join_params(uo.init_params(3, scale=300.0)).astype(np.float32)
    a, sa = run_learned(scene, params, k, True)
    s, ss = run_learned(scene, params, k, False)
    assert sa["fallback"] == ss["fallback"] == (net == "fallback")
    for key in s:
        assert np.array_equal(a[key], s[key]), f"k={k} {net}: {key}"
