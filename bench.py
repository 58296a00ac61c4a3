"""Throughput of the stable-fluids 3D step on one MI355X.

Workload (default): 256^3 cells, h = 1/256, dt = 0.1, nu = kappa = 1e-4, 40 diffusion and
80 pressure iterations (red-black Gauss-Seidel), free-slip walls, plume sphere of radius
16 at (128, 24, 128) with source 100 and upward force 5, initial velocity = seeded N(0,1)
noise smoothed by a 5^3 box filter. One step = one hipGraph launch of the whole step.

Printed JSON line: steps/s (value), Gcell-updates/s, the dominant kernel's algorithmic
bandwidth against the 8 TB/s HBM peak (timed live with HIP events on the context stream,
beside the average of the committed kernel trace of the same command, and again at 512^3
where the pass's working set no longer fits the 256 MB last-level cache), the Jacobi
pass bandwidth (256^3 and 512^3), the in-graph cost of one more smoother pass (RBGS and
Jacobi, pressure and diffusion, from step times at two pass counts), the learned-vs-plain
speedup at equal post-projection
divergence L2 (warm-start network + K passes against 80 plain passes, from common
states of the 256^3 plume; needs the trained parameters under stable-fluids_amd/weights)
and the host-loop baseline (the oracle on this host's cores, bounded samples).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid N] [--smoother rbgs|jacobi]
                  [--storage f32|f16|bf16|f32-refined] [--interval M]
                  [--no-cpu-baseline] [--no-learned] [--no-512] [--no-in-graph] [--stage-repeat R]
  python bench.py --learned-step K [--steps S] [--warmup W]   (the learned step alone, for traces)
  python bench.py --batch [--tuning-cache FILE]                (64 uneven simulations, below)

Precision study (--precision): the plume at 512^3 (centre (256,48,256), radius 32,
D=20, P=40, zero initial velocity), 50 steps with FP32, FP16 and BF16 iterate storage
(refinement cycles with an FP32 defect every M=10 passes) and with the same cycles on
FP32 storage. Prints one JSON line with the per-field max and L2 error of each run
against the plain FP32 run, the final divergence L2 (FP64 reduction on the GPU) and
ms/step of each.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import __graft_entry__ as entry  # noqa: E402

HBM_PEAK_GBS = 8000.0
# the dominant kernel: a middle velocity-diffusion RBGS pass on the black-cell buffers
# (diffusion solve, half-buffer in, half-buffer out, 128 x 16 tiles)
VELOCITY_PASS_KERNEL = "k_rbgs_half<0, 2, 0, 64, 8, 2>"
TRACE_FILE = os.path.join(ROOT, "profiles", "stage_kernels_plume256_kernel_trace.json")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_plume256_kernels.json")
STORAGE = {"f32": 0, "f16": 1, "bf16": 2, "f32-refined": 3}
METRIC = "steps/s at 256³; Jacobi HBM GB/s; learned-vs-plain speedup at equal div L2"


def step_bytes_per_cell(d_iters, p_iters, n_scalars=1, n_forces=1, alpha_pos=True, alpha_s_pos=True, dissipation=False):
    """Algorithmic bytes per cell of one step as the library's kernel set runs it on an
    even-nx grid (each array read once and written once per kernel). RBGS solves run on
    the black-cell passes: the first reads b and the start iterate (natural, 4 B, or the
    zero guess, none) and writes the black cells (2 B), the middle ones read b and the
    black cells and write the black cells (8 B), the last writes the natural field (10 B)."""
    def solve(k, zero):
        if k <= 0:
            return 0
        first = 4 + (0 if zero else 4)
        if k == 1:
            return first + 4
        return first + 2 + 8 * (k - 2) + 10
    b = 36 * n_forces + 24                                      # add force, fused 3-component advection
    b += 3 * solve(d_iters, False) if alpha_pos else 24         # velocity diffusion (or copy)
    b += 16 + (solve(p_iters, True) if p_iters > 0 else 4) + 36  # divergence, pressure, project+div+|v|
    for _ in range(n_scalars):
        b += 12 + 20                                            # source add, advection
        b += solve(d_iters, False) if alpha_s_pos else (8 if dissipation else 0)
    return b


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--smoother", choices=("rbgs", "jacobi"), default="rbgs")
    ap.add_argument("--stage-repeat", type=int, default=40)
    ap.add_argument("--storage", choices=tuple(STORAGE), default="f32")
    ap.add_argument("--interval", type=int, default=10, help="passes per refinement cycle (16-bit storage)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain projection comparison")
    ap.add_argument("--no-512", action="store_true", help="skip the 512^3 smoother-pass leg of the roofline")
    ap.add_argument("--no-in-graph", action="store_true", help="skip the in-graph marginal pass costs")
    ap.add_argument("--precision", action="store_true", help="run the 512^3 storage precision study instead")
    ap.add_argument("--fp64", action="store_true",
                    help="with --precision: the well-conditioned variant (force 0.5, 5 steps) against the FP64 oracle")
    ap.add_argument("--batch", action="store_true", help="run the 64-simulation batched study instead")
    ap.add_argument("--tuning-cache", default=None, help="autotuner cache file for --batch")
    ap.add_argument("--learned-step", type=int, default=None, metavar="K",
                    help="time only the learned-warm-start step (network guess + K RBGS passes, P = 80) on the 256^3 "
                         "plume from its step-10 state, --steps steps after --warmup (for kernel traces)")
    ap.add_argument("--stages-only", action="store_true",
                    help="only the stage-hook timings of the default workload (no step, no graph launch: for kernel "
                         "traces of the smoother passes)")
    ap.add_argument("--learned-guess-stage", type=int, default=None, metavar="R",
                    help="enqueue the learned projection stage (K = 0: the network guess, its residual test and the "
                         "projection) R times through the stage hook at 256^3, with no graph launch (for kernel traces)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the stable-fluids step here runs one simulation per GPU; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    pkg = entry.load_package()
    if args.precision:
        precision_study(pkg, args)
        return
    if args.batch:
        batch_study(pkg, args)
        return
    if args.learned_step is not None:
        learned_step_study(pkg, args)
        return
    if args.learned_guess_stage is not None:
        learned_guess_stage(pkg, args)
        return
    if args.stages_only:
        stages_only(pkg, args)
        return
    abi, hip, scenes = pkg.abi, pkg.hip, pkg.scenes
    n = args.grid
    scene = scenes.plume_scene(n, diffuse_iterations=40, pressure_iterations=80, noise_seed=0)
    smoother = abi.SMOOTHER_JACOBI if args.smoother == "jacobi" else abi.SMOOTHER_RBGS
    opts = pkg.ext_options(smoother=smoother, iterate_storage=STORAGE[args.storage], defect_correction_interval=args.interval)
    ctx = pkg.Context(scene.desc, options=opts)
    scenes.seed_context(ctx, scene)
    stream = ctx.stream
    import ctypes as C
    hs = C.c_void_p(stream)

    # warmup, then K graph launches bracketed by device synchronisation
    ctx.step(args.warmup)
    hip.synchronize()
    ev0, ev1 = hip.Event(), hip.Event()
    t0 = time.perf_counter()
    hip.check(hip.lib().hipEventRecord(ev0.handle, hs), "hipEventRecord")
    ctx.step(args.steps)
    hip.check(hip.lib().hipEventRecord(ev1.handle, hs), "hipEventRecord")
    hip.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev1.elapsed_ms(ev0)
    norms = ctx.norms()
    ms_per_step = 1e3 * wall / args.steps
    steps_per_s = args.steps / wall
    cells = scene.desc.cells
    bpc = step_bytes_per_cell(40, 80)

    # dominant kernel: a middle velocity-diffusion smoother pass (3 * 38 of them per step),
    # timed live on the context stream with HIP events
    stages = timed_stages(ctx, args.stage_repeat)
    dom = stages["velocity_pass"]
    tiling = ctx.smoother_tiling()
    ctx.destroy()
    default_run = n == 256 and args.smoother == "rbgs" and args.storage == "f32"
    # the same passes at 512^3 (working set 4x the last-level cache)
    big = None
    if default_run and not args.no_512:
        big_scene = scenes.plume_scene(512, diffuse_iterations=40, pressure_iterations=80)
        with pkg.Context(big_scene.desc) as big_ctx:
            scenes.seed_context(big_ctx, big_scene)
            big = timed_stages(big_ctx, max(4, args.stage_repeat // 4), names=("velocity_pass", "jacobi_pass"))

    # in-graph cost of one more smoother pass, per smoother and solve (no profiler)
    in_graph = None
    if default_run and not args.no_in_graph:
        in_graph = in_graph_pass_costs(pkg, n)

    # the committed kernel trace (rocprofv3 --kernel-trace --stats of the same stage launches,
    # bench.py --stages-only) and the committed FETCH_SIZE / WRITE_SIZE counts of the kernel
    # (copy-kernel calibrated)
    trace = traffic = None
    if default_run and os.path.exists(TRACE_FILE):
        trace = json.load(open(TRACE_FILE))["kernels"].get(VELOCITY_PASS_KERNEL)
    if default_run and os.path.exists(TRAFFIC_FILE):
        rec = json.load(open(TRAFFIC_FILE))["kernels"].get(VELOCITY_PASS_KERNEL)
        traffic = rec["traffic_bytes"] if rec else None

    # learned warm start vs plain pressure passes at equal post-projection divergence L2
    learned = None
    if not args.no_learned and n == 256 and args.storage == "f32":
        ev = pkg.warm_eval.evaluate(n=256)
        if ev is not None:
            learned = {
                "workload": "plume 256^3, centre (128,24,128) radius 16, zero initial velocity, D=20; one step from "
                            "checkpoints at steps 10/20/30; network on the 32^3 restriction",
                "plain_passes": ev["plain_passes"],
                "k_equal_rbgs": ev["k_equal_rbgs"], "k_equal_jacobi": ev["k_equal_jacobi"],
                "speedup_vs_rbgs": ev["speedup_vs_rbgs"], "speedup_vs_jacobi": ev["speedup_vs_jacobi"],
                "pressure_stage_speedup_vs_rbgs": ev.get("pressure_stage_speedup_vs_rbgs"),
                "pressure_stage_speedup_vs_jacobi": ev.get("pressure_stage_speedup_vs_jacobi"),
                "pressure_stage_ms": ev["pressure_stage_ms"],
                "projection_stage_ms": ev["projection_stage_ms"],
                "projection_stage_speedup_vs_rbgs": ev.get("projection_stage_speedup_vs_rbgs"),
                "projection_stage_speedup_vs_jacobi": ev.get("projection_stage_speedup_vs_jacobi"),
                "stage_timing": "each arm's pressure solve (and whole projection) enqueued alone through the stage hooks "
                                "after one step from the step-30 state, HIP events on the context stream",
                "plain_rbgs_div_l2": ev["plain_rbgs_div_l2"], "plain_jacobi_div_l2": ev["plain_jacobi_div_l2"],
                "learned_div_l2_by_k": {str(c["k"]): c["div_l2"] for c in ev["learned_curve"]},
                "fallbacks": sum(sum(c["fallback"]) for c in ev["learned_curve"]),
                "steps_per_s": {k: v for k, v in ev.items() if k.endswith("steps_per_s")},
            }

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(pkg, scene, steps_per_s)

    out = {
        "metric": METRIC,
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "gpu_ms_per_step": round(gpu_ms / args.steps, 4),
        "gcell_updates_per_s": round(cells * steps_per_s / 1e9, 4),
        "step_hbm_gbs": round(bpc * cells * steps_per_s / 1e9, 1),
        "step_bytes_per_cell": bpc,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plume {n}^3, seeded smoothed-noise initial velocity, D=40 P=80 {args.smoother}, hipGraph step",
                   "grid": [n, n, n], "diffuse_iterations": 40, "pressure_iterations": 80, "smoother": args.smoother,
                   "iterate_storage": args.storage, "refinement_interval": args.interval if args.storage != "f32" else None,
                   "dt": 0.1, "viscosity": 1e-4, "diffusion": 1e-4, "smoother_tiling": list(tiling)},
        "roofline": roofline(dom, trace, traffic, big),
        "jacobi_hbm_gbs": stages["jacobi_pass"]["gbs"],
        "jacobi_hbm_gbs_512": big["jacobi_pass"]["gbs"] if big else None,
        "in_graph_pass": in_graph,
        "learned_vs_plain_speedup": learned["speedup_vs_rbgs"] if learned else None,
        "learned_vs_plain": learned,
        "stages": stages,
        "final_divergence_l2": norms["divergence_l2"],
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def timed_stages(ctx, repeat, names=None):
    """Mean GPU time and algorithmic bandwidth of single stage launches (stage hooks)."""
    stage_list = (("velocity_pass", 1), ("pressure_pass", 0), ("jacobi_pass", 4), ("advect", 2), ("project", 3), ("copy", 5))
    out = {}
    for name, stage in stage_list:
        if names is not None and name not in names:
            continue
        ms = ctx.time_stage(stage, repeat)
        by = ctx.stage_bytes(stage)
        out[name] = {"us": round(ms * 1e3, 3), "bytes": by, "gbs": round(by / (ms * 1e-3) / 1e9, 1)}
    return out


def in_graph_pass_costs(pkg, n, reps=3, steps=10):
    """Marginal cost of one smoother pass inside the captured step, without a profiler: the
    step's GPU time (best of `reps` runs of `steps` graph launches) at P = 160 minus P = 80
    over the 80 extra pressure passes, and at D = 80 minus D = 40 over the 160 extra
    diffusion passes (three velocity components and one scalar). It includes each pass's
    launch as a graph node."""
    abi, hip, scenes = pkg.abi, pkg.hip, pkg.scenes
    alg = {"rbgs": 8 * n ** 3, "jacobi": 12 * n ** 3}  # algorithmic bytes of a middle pass
    out = {"method": "step time at P=160 minus P=80 (per pressure pass) and at D=80 minus D=40 (per diffusion pass), "
                     f"best of {reps} x {steps} graph launches; includes the per-node launch"}
    for name, sm in (("rbgs", abi.SMOOTHER_RBGS), ("jacobi", abi.SMOOTHER_JACOBI)):
        ms = {}
        for d, p in ((40, 80), (40, 160), (80, 80)):
            scene = scenes.plume_scene(n, diffuse_iterations=d, pressure_iterations=p, noise_seed=0)
            with pkg.Context(scene.desc, options=pkg.ext_options(smoother=sm)) as ctx:
                scenes.seed_context(ctx, scene)
                ctx.step(3)
                hip.synchronize()
                best = 1e9
                for _ in range(reps):
                    t0 = time.perf_counter()
                    ctx.step(steps)
                    hip.synchronize()
                    best = min(best, (time.perf_counter() - t0) / steps)
                ms[(d, p)] = best * 1e3
        pres = (ms[(40, 160)] - ms[(40, 80)]) / 80 * 1e3
        diff = (ms[(80, 80)] - ms[(40, 80)]) / 160 * 1e3
        out[name] = {"step_ms": round(ms[(40, 80)], 4), "pressure_pass_us": round(pres, 3), "diffusion_pass_us": round(diff, 3),
                     "bytes_per_pass": alg[name],
                     "pressure_frac": round(alg[name] / (pres * 1e-6) / 1e9 / HBM_PEAK_GBS, 4),
                     "diffusion_frac": round(alg[name] / (diff * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)}
    return out


def roofline(dom, trace, traffic, big):
    """The dominant kernel against the HBM peak. achieved = its algorithmic bytes (8 B per
    cell: b and the black cells read, the black cells written) over its live mean launch
    time; trace_* are the same from the committed kernel trace of these stage launches
    (bench.py --stages-only)."""
    out = {"bound": "hbm", "kernel": f"velocity diffusion RBGS middle pass ({VELOCITY_PASS_KERNEL})", "achieved": dom["gbs"],
           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom["gbs"] / HBM_PEAK_GBS, 4), "traffic": traffic,
           "bytes_per_launch": dom["bytes"], "avg_us": dom["us"],
           "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT) if traffic else None,
           "note": "at 256^3 the pass's 128 MB working set fits the 256 MB last-level cache, so part of its reads "
                   "come from there; frac_512 is the same pass where it does not"}
    if trace:
        t_gbs = dom["bytes"] / (trace["avg_us"] * 1e-6) / 1e9
        out.update({"trace_avg_us": trace["avg_us"], "trace_frac": round(t_gbs / HBM_PEAK_GBS, 4),
                    "trace_source": os.path.relpath(TRACE_FILE, ROOT)})
    if big:
        v = big["velocity_pass"]
        out.update({"avg_us_512": v["us"], "achieved_512": v["gbs"], "frac_512": round(v["gbs"] / HBM_PEAK_GBS, 4)})
    return out


def cpu_baseline(pkg, scene, gpu_steps_per_s):
    """The oracle on this host's cores: the 256^3 workload (FP64 and FP32, one step each,
    all threads) and the 64^3 plume of 20/40 passes (FP64, 100 steps, all threads; FP64
    and FP32, 3 steps, one thread)."""
    abi, scenes = pkg.abi, pkg.scenes
    oracle = entry.load_oracle()
    threads = oracle.threads(abi)
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count()
    small = scenes.plume_scene(64)

    def run(sc, precision, n_threads, steps):
        oracle.set_threads(abi, n_threads)
        sim = oracle.OracleSim(abi, sc.desc, precision=precision)
        try:
            scenes.seed_oracle(sim, sc)
            t = time.perf_counter()
            sim.step(steps)
            dt = time.perf_counter() - t
        finally:
            sim.destroy()
            oracle.set_threads(abi, threads)
        return {"steps": steps, "seconds": round(dt, 3), "steps_per_s": round(steps / dt, 5),
                "gcell_updates_per_s": round(sc.desc.cells * steps / dt / 1e9, 6), "threads": n_threads}

    runs = {"plume256_fp64_all": run(scene, 64, threads, 1), "plume256_fp32_all": run(scene, 32, threads, 1),
            "plume64_fp64_all_100_steps": run(small, 64, threads, 100), "plume64_fp64_1_thread": run(small, 64, 1, 3),
            "plume64_fp32_1_thread": run(small, 32, 1, 3)}
    main_run = runs["plume256_fp64_all"]
    total = sum(r["seconds"] for r in runs.values())
    return {"value": main_run["steps_per_s"], "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"1 step of the same 256^3 workload (FP64 host loops, {threads} OpenMP threads); beside it FP32 and "
                      f"the 64^3 plume (D=20, P=40) with 1 and {threads} threads, {total:.1f} s in all",
            "nproc": nproc, "omp_max_threads": threads, "runs": runs,
            "gpu_over_cpu": round(gpu_steps_per_s / main_run["steps_per_s"], 1)}


def learned_step_study(pkg, args):
    """Steps/s of the learned-warm-start step alone (the evaluation's timing setup)."""
    we, abi, scenes = pkg.warm_eval, pkg.abi, pkg.scenes
    params = we.load_model()
    if params is None:
        print("no trained warm-start parameters under stable-fluids_amd/weights", file=sys.stderr)
        sys.exit(2)
    scene = scenes.plume_scene(256, diffuse_iterations=20, pressure_iterations=80)
    state = we.checkpoints(scene, (10,))[-1]
    opts = pkg.ext_options(smoother=abi.SMOOTHER_RBGS, pressure_warm_start=abi.WARM_START_LEARNED,
                           warm_start_iterations=args.learned_step, warm_start_coarsening=3)
    sps = we.steps_per_second(scene, state, opts, params, steps=args.steps, warmup=args.warmup)
    print(json.dumps({"workload": "learned warm-start step, plume 256^3 from its step-10 state, D=20 P=80, "
                                  "network on the 32^3 restriction", "k": args.learned_step, "steps": args.steps,
                      "steps_per_s": round(sps, 3)}), flush=True)


def stages_only(pkg, args):
    """The `stages` object of the default run alone: each stage kernel enqueued through the
    stage hooks on the seeded plume, with no step and no graph launch before it. One
    advection and one projection enqueued first give the smoothers nonzero right-hand sides
    (the advected velocity, the divergence); the seeded velocity is white noise, so the
    advection's own time is not the plume's."""
    abi, scenes = pkg.abi, pkg.scenes
    scene = scenes.plume_scene(args.grid, diffuse_iterations=40, pressure_iterations=80, noise_seed=0)
    with pkg.Context(scene.desc) as ctx:
        scenes.seed_context(ctx, scene)
        ctx.run_stage(abi.STAGE_ADVECT, 1)
        ctx.run_stage(abi.STAGE_PROJECTION, 1)
        stages = timed_stages(ctx, args.stage_repeat)
        tiling = ctx.smoother_tiling()
    print(json.dumps({"workload": f"plume {args.grid}^3, seeded smoothed-noise velocity, D=40 P=80 RBGS, stage hooks only",
                      "stage_repeat": args.stage_repeat, "stages": stages, "tiling": tiling}), flush=True)


def learned_guess_stage(pkg, args):
    """GPU ms of the learned projection stage with K = 0 at 256^3 (divergence, restriction,
    network, prolongation, residual test, the exiting fallback launch, projection), enqueued
    eagerly through the stage hook: no step and no graph launch precede it. The context
    holds the seeded plume at rest, so the divergence is zero and the all-zero network's
    guess is accepted; every kernel still runs its whole grid."""
    import numpy as np
    abi, scenes = pkg.abi, pkg.scenes
    scene = scenes.plume_scene(256, diffuse_iterations=20, pressure_iterations=80)
    opts = pkg.ext_options(smoother=abi.SMOOTHER_RBGS, pressure_warm_start=abi.WARM_START_LEARNED,
                           warm_start_iterations=0, warm_start_coarsening=3)
    with pkg.Context(scene.desc, options=opts) as ctx:
        ctx.set_pressure_model(np.zeros(pkg.unet.param_count(), np.float32))
        scenes.seed_context(ctx, scene)
        ms = ctx.time_stage(abi.STAGE_PROJECTION, args.learned_guess_stage)
        stats = ctx.warm_start_stats()
    print(json.dumps({"workload": "learned projection stage, K = 0, plume 256^3 at rest, network on the 32^3 "
                                  "restriction (all-zero parameters)", "repeat": args.learned_guess_stage,
                      "stage_ms": round(ms, 4), "fallback": stats["fallback"]}), flush=True)


def precision_study(pkg, args):
    """Per-field error of 16-bit iterate storage against FP32 storage after K steps (and,
    with --fp64, of every storage against the FP64 oracle on a well-conditioned scene)."""
    import numpy as np
    abi, hip, scenes = pkg.abi, pkg.hip, pkg.scenes
    n = args.grid if args.grid != 256 else 512
    steps = args.steps if args.steps != 200 else (5 if args.fp64 else 50)
    force = 0.5 if args.fp64 else 5.0
    scene = scenes.plume_scene(n, diffuse_iterations=20, pressure_iterations=40, force=force)
    smoother = abi.SMOOTHER_JACOBI if args.smoother == "jacobi" else abi.SMOOTHER_RBGS
    kinds = {"pressure": abi.VIEW_PRESSURE, "divergence": abi.VIEW_DIVERGENCE, "speed": abi.VIEW_VELOCITY_MAGNITUDE}

    def run(storage):
        opts = pkg.ext_options(smoother=smoother, iterate_storage=STORAGE[storage], defect_correction_interval=args.interval)
        ctx = pkg.Context(scene.desc, options=opts)
        try:
            scenes.seed_context(ctx, scene)
            ctx.step(1)
            hip.synchronize()
            # restart from the seeded state so every run takes exactly `steps` steps
            ctx.set_flow_velocity(None, None, None)
            ctx.update_scalar_field(ctx.scalar_handles[0], None)
            hip.synchronize()
            t = time.perf_counter()
            ctx.step(steps)
            hip.synchronize()
            ms = 1e3 * (time.perf_counter() - t) / steps
            fields = {k: ctx.read(v) for k, v in kinds.items()}
            fields.update(zip(("vx", "vy", "vz"), ctx.read(abi.VIEW_VELOCITY)))
            fields["density"] = ctx.read(abi.VIEW_SCALAR_DATA, 1)
            return fields, ctx.norms(), ms
        finally:
            ctx.destroy()

    def errors(got, ref):
        out = {}
        for k, r in ref.items():
            r64, g64 = r.astype(np.float64), got[k].astype(np.float64)
            scale, l2 = float(np.max(np.abs(r64))), float(np.sqrt(np.sum(r64 * r64)))
            d = g64 - r64
            out[k] = {"max_rel": float(np.max(np.abs(d))) / max(scale, 1e-300),
                      "l2_rel": float(np.sqrt(np.sum(d * d))) / max(l2, 1e-300)}
            del r64, g64, d
        return out

    o64 = None
    if args.fp64:
        # the FP64 oracle on this host's cores, one step at a time (progress on stdout)
        oracle = entry.load_oracle()
        sim = oracle.OracleSim(abi, scene.desc, precision=64, smoother=oracle.JACOBI if args.smoother == "jacobi" else oracle.RBGS)
        try:
            scenes.seed_oracle(sim, scene)
            t = time.perf_counter()
            for i in range(steps):
                sim.step(1)
                print(f"fp64 oracle step {i + 1}/{steps}: {time.perf_counter() - t:.1f} s", flush=True)
            o64 = {k: sim.read(v) for k, v in kinds.items()}
            o64.update(zip(("vx", "vy", "vz"), sim.read(abi.VIEW_VELOCITY)))
            o64["density"] = sim.read(abi.VIEW_SCALAR_DATA, 1)
        finally:
            sim.destroy()

    ref, ref_norms, ref_ms = run("f32")
    runs = {"f32": {"ms_per_step": round(ref_ms, 3), "final_divergence_l2": ref_norms["divergence_l2"]}}
    if o64 is not None:
        runs["f32"]["error_vs_fp64_oracle"] = errors(ref, o64)
    # f32-refined runs the same cycles with FP32 storage: its difference from the plain
    # run is rounding order alone, amplified by the flow, which bounds what storage can show
    for storage in ("f32-refined", "f16", "bf16"):
        got, norms, ms = run(storage)
        runs[storage] = {"ms_per_step": round(ms, 3), "speedup_vs_f32": round(ref_ms / ms, 3),
                         "final_divergence_l2": norms["divergence_l2"], "error_vs_f32_storage": errors(got, ref)}
        if o64 is not None:
            runs[storage]["error_vs_fp64_oracle"] = errors(got, o64)
        del got
    print(json.dumps({"study": "iterate storage precision", "grid": [n, n, n], "steps": steps, "smoother": args.smoother,
                      "refinement_interval": args.interval, "diffuse_iterations": 20, "pressure_iterations": 40,
                      "workload": f"plume {n}^3, centre ({n // 2},{3 * n // 32},{n // 2}) radius {n // 16}, force {force}, "
                                  f"zero initial velocity", "fp64_oracle": o64 is not None,
                      "runs": runs}), flush=True)


def batch_study(pkg, args):
    """Aggregate throughput of 64 uneven simulations: one-by-one vs one batched launch per stage."""
    hip = pkg.hip
    steps = args.steps if args.steps != 200 else 10
    scenes_ = pkg.scenes.batch_plume_scenes()
    cells = sum(sc.desc.cells for sc in scenes_)

    def run(make_stream, batched, tune=False):
        shared = hip.Stream() if make_stream else None
        ctxs = []
        for sc in scenes_:
            ctx = pkg.Context(sc.desc, stream=shared)
            pkg.scenes.seed_context(ctx, sc)
            ctxs.append(ctx)
        batch = pkg.Batch(ctxs) if batched else None
        info = None
        if batch is not None and tune:
            tuned = batch.autotune(args.tuning_cache, repeats=20)
            plans = {}
            for i, sc in enumerate(scenes_):
                p = batch.info(i)
                plans["x".join(map(str, reversed(sc.desc.shape)))] = {"tile": p["tile"], "z_chunk": p["z_chunk"]}
            info = {"tuned_shapes": tuned, "plans": plans}
        step = batch.step if batch else (lambda: [c.step(1) for c in ctxs])
        step()
        hip.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            step()
        hip.synchronize()
        dt = time.perf_counter() - t
        if batch is not None:
            info = dict(info or {}, launches_per_step=batch.info(0)["launches_per_step"])
            batch.destroy()
        for c in ctxs:
            c.destroy()
        if shared is not None:
            shared.destroy()
        return steps / dt, info

    res = {}
    for name, shared, batched, tune in (("one_by_one_shared_stream", True, False, False), ("one_by_one_own_streams", False, False, False),
                                        ("batched", False, True, False), ("batched_autotuned", False, True, True)):
        sps, info = run(shared, batched, tune)
        res[name] = {"batch_steps_per_s": round(sps, 3), "gcell_updates_per_s": round(cells * sps / 1e9, 4)}
        if info:
            res[name].update(info)
    print(json.dumps({"study": "batched uneven simulations", "simulations": len(scenes_), "total_cells": cells,
                      "diffuse_iterations": 40, "pressure_iterations": 80, "steps": steps,
                      "sizes": sorted({"x".join(map(str, reversed(sc.desc.shape))) for sc in scenes_}),
                      "results": res, "speedup_batched_vs_shared_stream": round(res["batched_autotuned"]["batch_steps_per_s"] / res["one_by_one_shared_stream"]["batch_steps_per_s"], 3),
                      "speedup_batched_vs_own_streams": round(res["batched_autotuned"]["batch_steps_per_s"] / res["one_by_one_own_streams"]["batch_steps_per_s"], 3)}), flush=True)


if __name__ == "__main__":
    main()
