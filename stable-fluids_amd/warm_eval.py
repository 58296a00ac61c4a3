"""Learned-vs-plain pressure projection at equal post-projection divergence L2.

From common simulation states (velocity and density checkpointed from a plain run), one
step is taken with the plain solver (pressure from zero, P passes) and with the learned
warm start (network guess + K passes) for a range of K. The smallest K whose
divergence L2 matches the plain P-pass value at every checkpoint is the equal-quality
setting. Both settings are then timed on the GPU: whole steps/s, and the step's own
pressure solve and projection (divergence + solve + gradient subtraction), each
enqueued alone through the stage hooks from the same state and timed with HIP events on
the context stream -- the summed time of that arm's own kernels.
"""
import os
import time

import numpy as np

from . import abi, hip, runtime, scenes

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "unet_warm_start_plume.f32")


def load_model(path=WEIGHTS):
    if not os.path.exists(path):
        return None
    return np.fromfile(path, dtype="<f4").astype(np.float32)


def checkpoints(scene, at_steps, options=None):
    """[(step, (vx, vy, vz), density)] of a plain run of `scene`."""
    out = []
    with runtime.Context(scene.desc, options=options) as ctx:
        scenes.seed_context(ctx, scene)
        done = 0
        for s in sorted(at_steps):
            ctx.step(s - done)
            done = s
            out.append((s, ctx.read(abi.VIEW_VELOCITY), ctx.read(abi.VIEW_SCALAR_DATA, 1)))
    return out


def _restore(ctx, scene, state):
    scenes.seed_context(ctx, scene)
    _, vel, rho = state
    ctx.set_flow_velocity(*(ctx.upload(c) for c in vel))
    ctx.update_scalar_field(ctx.scalar_handles[0], ctx.upload(rho))


def one_step_div(scene, states, options, params=None):
    """Post-projection divergence L2 after one step from each state (+ warm-start stats)."""
    res, stats = [], []
    with runtime.Context(scene.desc, options=options) as ctx:
        if params is not None:
            ctx.set_pressure_model(params)
        for st in states:
            _restore(ctx, scene, st)
            ctx.step(1)
            res.append(ctx.norms()["divergence_l2"])
            if params is not None:
                stats.append(ctx.warm_start_stats())
    return res, stats


def steps_per_second(scene, state, options, params=None, steps=20, warmup=3):
    with runtime.Context(scene.desc, options=options) as ctx:
        if params is not None:
            ctx.set_pressure_model(params)
        _restore(ctx, scene, state)
        ctx.step(warmup)
        hip.synchronize()
        t = time.perf_counter()
        ctx.step(steps)
        hip.synchronize()
        return steps / (time.perf_counter() - t)


def stage_ms(scene, state, options, params=None, repeat=10):
    """GPU ms of the step's pressure solve and of its whole projection, after one step
    from `state` (the hooks then see that step's divergence and diffused velocity)."""
    with runtime.Context(scene.desc, options=options) as ctx:
        if params is not None:
            ctx.set_pressure_model(params)
        _restore(ctx, scene, state)
        ctx.step(1)
        return {"pressure_solve": ctx.time_stage(abi.STAGE_PRESSURE_SOLVE, repeat),
                "projection": ctx.time_stage(abi.STAGE_PROJECTION, repeat)}


def evaluate(n=256, plain_passes=80, k_list=(2, 4, 6, 8, 10, 12, 13, 14, 15, 16, 17, 18, 19, 20, 24, 32, 40, 56, 80), at_steps=(10, 20, 30),
             coarsening=3, params=None, time_steps=60):
    """Equal-quality comparison on the plume at n^3 (centre (n/2, 3n/32, n/2), radius n/16,
    zero initial velocity, D = 20). Returns a dict (None when no model is available)."""
    params = load_model() if params is None else params
    if params is None:
        return None
    scene = scenes.plume_scene(n, diffuse_iterations=20, pressure_iterations=plain_passes)
    states = checkpoints(scene, at_steps)
    rbgs = runtime.ext_options(smoother=abi.SMOOTHER_RBGS)
    jac = runtime.ext_options(smoother=abi.SMOOTHER_JACOBI)
    plain_div, _ = one_step_div(scene, states, rbgs)
    jac_div, _ = one_step_div(scene, states, jac)
    curve = []
    k_rbgs = k_jac = None
    for k in k_list:
        if k > plain_passes:
            break
        opts = runtime.ext_options(smoother=abi.SMOOTHER_RBGS, pressure_warm_start=abi.WARM_START_LEARNED,
                                   warm_start_iterations=k, warm_start_coarsening=coarsening)
        d, st = one_step_div(scene, states, opts, params)
        curve.append({"k": k, "div_l2": d, "fallback": [s["fallback"] for s in st],
                      "guess_over_zero_residual": [s["guess_residual"] / max(s["zero_residual"], 1e-300) for s in st]})
        if k_rbgs is None and all(a <= b for a, b in zip(d, plain_div)):
            k_rbgs = k
        if k_jac is None and all(a <= b for a, b in zip(d, jac_div)):
            k_jac = k
    out = {"grid": n, "plain_passes": plain_passes, "checkpoint_steps": list(at_steps), "coarsening": coarsening,
           "plain_rbgs_div_l2": plain_div, "plain_jacobi_div_l2": jac_div, "learned_curve": curve,
           "k_equal_rbgs": k_rbgs, "k_equal_jacobi": k_jac}
    base_state = states[-1]
    plain_sps = steps_per_second(scene, base_state, rbgs, steps=time_steps)
    jac_sps = steps_per_second(scene, base_state, jac, steps=time_steps)
    out["plain_rbgs_steps_per_s"] = plain_sps
    out["plain_jacobi_steps_per_s"] = jac_sps
    st_rbgs = stage_ms(scene, base_state, rbgs)
    st_jac = stage_ms(scene, base_state, jac)
    out["pressure_stage_ms"] = {"plain_rbgs": st_rbgs["pressure_solve"], "plain_jacobi": st_jac["pressure_solve"]}
    out["projection_stage_ms"] = {"plain_rbgs": st_rbgs["projection"], "plain_jacobi": st_jac["projection"]}
    for tag, k, base, st_base in (("rbgs", k_rbgs, plain_sps, st_rbgs), ("jacobi", k_jac, jac_sps, st_jac)):
        if k is None:
            out[f"speedup_vs_{tag}"] = None
            continue
        # the learned context (P = plain_passes, of which the first P - K run only on fallback)
        opts = runtime.ext_options(smoother=abi.SMOOTHER_RBGS, pressure_warm_start=abi.WARM_START_LEARNED,
                                   warm_start_iterations=k, warm_start_coarsening=coarsening)
        sps = steps_per_second(scene, base_state, opts, params, steps=time_steps)
        st = stage_ms(scene, base_state, opts, params)
        out[f"learned_k{k}_steps_per_s"] = sps
        out[f"speedup_vs_{tag}"] = sps / base
        out["pressure_stage_ms"][f"learned_k{k}"] = st["pressure_solve"]
        out["projection_stage_ms"][f"learned_k{k}"] = st["projection"]
        out[f"pressure_stage_speedup_vs_{tag}"] = st_base["pressure_solve"] / st["pressure_solve"]
        out[f"projection_stage_speedup_vs_{tag}"] = st_base["projection"] / st["projection"]
    return out
