"""Device-side animated forcing (stable_fluids_ext_set_force_animation_cuda) on the
visualizer's plume scene (vulkan-app/scene_plume.cpp:241-335), which the reference drives
by recomputing force x/z on the host and uploading them before every step.

* the force the step graph computes equals the FP32 host restatement of the schedule
  (scenes.animated_force) at the counter's step, within the ulps of sin/cos;
* a context with the animation ends bit-identical to one whose forces are uploaded before
  every step with the same values: the animation kernel feeds the step exactly like
  update_vector_field;
* the counter starts at start_step and advances once per step; turning the animation off
  keeps the last force; an animated context cannot join a batch.
"""
import numpy as np
import pytest

from conftest import sf
from paritylib import gpu_fields

abi = sf.abi
scenes = sf.scenes
pytestmark = pytest.mark.gpu


def viewer(n=(32, 48, 32)):
    return scenes.viewer_plume_scene(*n, diffuse_iterations=8, pressure_iterations=16)


def animated_context(scene, masks, start):
    ctx = sf.Context(scene.desc)
    scenes.seed_context(ctx, scene)
    ctx.set_force_animation(ctx.vector_handles[0], abi.plume_scene_animation(start), *(ctx.upload(m) for m in masks))
    return ctx


def test_animated_force_matches_the_schedule():
    scene, masks = viewer()
    anim = abi.plume_scene_animation(3)
    with animated_context(scene, masks, 3) as ctx:
        assert ctx.force_animation_step() == 3
        for steps in (1, 4, 30):
            ctx.step(steps)
            n = ctx.force_animation_step()
            fx, fy, fz = ctx.read(abi.VIEW_VECTOR, ctx.vector_handles[0])
            ex, ez = scenes.animated_force(masks, n - 1, anim)
            scale = float(max(np.abs(ex).max(), np.abs(ez).max()))
            err = max(float(np.abs(fx - ex).max()), float(np.abs(fz - ez).max())) / scale
            print(f"step {n - 1}: max |force - host| / max|force| = {err:.2e}", flush=True)
            assert err <= 1e-6
            assert np.array_equal(fy, scene.force[1])
        assert ctx.force_animation_step() == 3 + 35


def test_animation_equals_uploaded_forces_bitwise():
    scene, masks = viewer()
    steps = 6
    with animated_context(scene, masks, 0) as a, sf.Context(scene.desc) as b:
        scenes.seed_context(b, scene)
        for _ in range(steps):
            a.step(1)
            fx, fy, fz = a.read(abi.VIEW_VECTOR, a.vector_handles[0])
            b.update_vector_field(b.vector_handles[0], b.upload(fx), b.upload(fy), b.upload(fz))
            b.step(1)
        ga, gb = gpu_fields(a, scene.desc), gpu_fields(b, scene.desc)
    for k in gb:
        assert np.array_equal(ga[k], gb[k]), k
    assert float(np.abs(ga["vx"]).max()) > 0.0  # the swirl drove a horizontal flow


def test_animation_off_keeps_the_last_force_and_blocks_batches():
    scene, masks = viewer()
    with animated_context(scene, masks, 0) as ctx:
        with pytest.raises(sf.StableFluidsError):
            sf.Batch([ctx])
        ctx.step(2)
        last = ctx.read(abi.VIEW_VECTOR, ctx.vector_handles[0])
        ctx.set_force_animation(ctx.vector_handles[0], None)
        ctx.step(2)
        now = ctx.read(abi.VIEW_VECTOR, ctx.vector_handles[0])
        for a, b in zip(last, now):
            assert np.array_equal(a, b)
        sf.Batch([ctx]).destroy()
