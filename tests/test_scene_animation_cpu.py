"""Host restatement of the visualizer plume's forcing masks and schedule
(scenes.viewer_plume_scene / animated_force; vulkan-app/scene_plume.cpp:241-289, :327-334),
the reference the GPU animation test compares against: the swirl mask is the unit
horizontal tangent times its radial weight, the weights are 1 at the centre and 0 from the
radius on, and at step 0 (phase 0) the schedule reduces to x = 1.65 swirl_x and
z = 0.85 drift + 1.65 swirl_z."""
import numpy as np

from conftest import sf

scenes = sf.scenes
abi = sf.abi


def test_masks_are_weighted_unit_tangents():
    scene, (drift, sx, sz) = scenes.viewer_plume_scene(32, 48, 32)
    f32 = np.float32
    h = f32(0.01)
    k, j, i = np.meshgrid(np.arange(32, dtype=f32), np.arange(48, dtype=f32), np.arange(32, dtype=f32), indexing="ij")
    dx, dz = (i + f32(0.5)) * h - f32(32) * h * f32(0.5), (k + f32(0.5)) * h - f32(32) * h * f32(0.5)
    mag = np.sqrt(sx.astype(np.float64) ** 2 + sz.astype(np.float64) ** 2)
    # perpendicular to the horizontal radius, counter-clockwise seen from +y
    assert np.max(np.abs(sx * dx + sz * dz)) <= 1e-7
    assert np.all(sz * dx - sx * dz >= -1e-9)
    # |tangent| = weight in (0, 1]; zero outside the swirl sphere
    assert mag.max() <= 1.0 and mag.max() > 0.9 and np.count_nonzero(mag) > 0
    assert drift.max() <= 1.0 and drift.min() >= 0.0 and scene.source.max() <= 32.0
    assert np.array_equal(scene.force[1] > 0, scene.source > 0)


def test_schedule_at_phase_zero_and_later():
    _, masks = scenes.viewer_plume_scene(32, 48, 32)
    drift, sx, sz = masks
    anim = abi.plume_scene_animation()
    fx, fz = scenes.animated_force(masks, 0, anim)
    f32 = np.float32
    assert np.array_equal(fx, (f32(0.0) * drift + f32(1.65) * sx).astype(f32))
    assert np.array_equal(fz, (f32(0.85) * drift + f32(1.65) * sz).astype(f32))
    # step 40: phase 1.8, coefficients against float64 evaluation
    fx, fz = scenes.animated_force(masks, 40, anim)
    p = 40 * 0.045
    ex = 1.15 * np.sin(p) * drift.astype(np.float64) + 1.65 * np.cos(p * 0.53) * sx.astype(np.float64)
    ez = 0.85 * np.cos(p * 0.71) * drift.astype(np.float64) + 1.65 * np.cos(p * 0.53) * sz.astype(np.float64)
    assert np.max(np.abs(fx - ex)) <= 1e-6 and np.max(np.abs(fz - ez)) <= 1e-6
