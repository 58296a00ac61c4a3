"""Trains the pressure warm-start network on pairs the solver itself generates.

Each pair comes from a running simulation at 128^3 (free-slip box, plume with random
centre, radius and force, optional smoothed-noise initial velocity, dt = 0.1,
nu = kappa = 1e-4, 20 diffusion / 40 pressure passes): after a step, the step's
pre-projection divergence is solved exactly (FP64 conjugate gradients to relative
residual 1e-6, stable_fluids_ext_training_pair_cuda) and the network learns
(div / sigma, wall mask) -> p* / (h^2 sigma) under a per-pair relative squared error
(weight 1 / mean(target^2)) of the mean-free prediction: in the closed box p* is defined
up to a constant (returned zero-mean), which neither the projection nor the guess test
sees. Forward, backward and Adam run on the
library's HIP kernels (stable-fluids-unet.h); one network (and its Adam moments) spans
all scenes.

The last layer is trained against target / S (S a power of two fixed from the first
pairs) and S is folded back into its weights when saving, so the saved network outputs
p* / (h^2 sigma) directly.

  python stable-fluids_amd/train_warm_start.py --pairs 20000 --max-seconds 900 --out <file>.f32

Writes <out> (raw little-endian float32 parameters, stable-fluids-unet.h layout) and
<out>.json (training record). Progress lines go to stdout.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(_HERE))

import __graft_entry__ as entry  # noqa: E402


def he_init(layers, seed):
    rng = np.random.default_rng(seed)
    out = []
    for ci, co in layers:
        out.append(rng.standard_normal(27 * co * ci) * math.sqrt(2.0 / (27 * ci)))
        out.append(np.zeros(co))
    return np.concatenate(out).astype(np.float32)


LAYERS = [(2, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 32), (32, 32), (32, 16), (16, 16), (16, 1)]


def last_layer_slice(count):
    w = 27 * 16 * 1
    return slice(count - w - 1, count)


def random_scene(pkg, n, rng, index):
    scenes, abi = pkg.scenes, pkg.abi
    desc = scenes.plume_desc(n, 20, 40)
    centre = (int(rng.integers(3 * n // 10, 7 * n // 10)), int(rng.integers(n // 16, n // 4)), int(rng.integers(3 * n // 10, 7 * n // 10)))
    radius = int(rng.integers(max(2, n // 32), n // 10 + 1))
    force = float(rng.uniform(1.0, 10.0))
    src, fx, fy, fz = scenes.plume_fields(desc.shape, centre, radius, 100.0, force)
    vel = scenes.noise_velocity(desc.shape, 1000 + index) if rng.random() < 0.5 else None
    if vel is not None:
        vel = tuple((c * np.float32(rng.uniform(0.2, 2.0))).astype(np.float32) for c in vel)
    return scenes.Scene(f"train{index}", desc, src, (fx, fy, fz), vel)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--pairs", type=int, default=20000)
    ap.add_argument("--steps-per-scene", type=int, default=40)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--lr-final", type=float, default=2e-5)
    ap.add_argument("--skip-steps", type=int, default=3, help="steps of each scene before its first pair")
    ap.add_argument("--max-seconds", type=float, default=900.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--init", default=None, help="start from these parameters (raw float32)")
    ap.add_argument("--coarsening", type=int, default=2, help="network grid = grid / 2^coarsening")
    ap.add_argument("--replay", type=int, default=64, help="pairs kept on the GPU for replay")
    ap.add_argument("--updates-per-pair", type=int, default=4, help="Adam steps per new pair (on replayed pairs)")
    ap.add_argument("--fixed-pairs", type=int, default=0, help="stop generating after this many pairs (overfit check)")
    ap.add_argument("--out", required=True)
    args = ap.parse_args()

    pkg = entry.load_package()
    hip, U = pkg.hip, pkg.unet
    n = args.grid
    count = U.param_count()
    rng = np.random.default_rng(args.seed)
    params = np.fromfile(args.init, dtype=np.float32) if args.init else he_init(LAYERS, args.seed)
    assert params.size == count

    nc = n >> args.coarsening
    cells = nc ** 3
    inp = hip.DeviceBuffer(cells * 8)
    tgt = hip.DeviceBuffer(cells * 4)
    R = max(1, args.replay)
    rep_inp = hip.DeviceBuffer(R * cells * 8)
    rep_tgt = hip.DeviceBuffer(R * cells * 4)
    rep_w = np.zeros(R)
    filled = 0
    updates = 0
    loss_d = hip.DeviceBuffer(4)
    t0 = time.perf_counter()
    last_print = t0
    record = {"grid": n, "pairs_requested": args.pairs, "losses": []}
    scale = None
    net = None
    ctx = None
    pair = 0
    scene_index = 0
    ema = None
    cg_its_total = 0
    while pair < args.pairs and time.perf_counter() - t0 < args.max_seconds:
        if ctx is not None:
            ctx.destroy()
        scene = random_scene(pkg, n, rng, scene_index)
        scene_index += 1
        ctx = pkg.Context(scene.desc)
        pkg.scenes.seed_context(ctx, scene)
        if args.skip_steps > 0 and (args.fixed_pairs <= 0 or pair < args.fixed_pairs):
            ctx.step(args.skip_steps)
        if net is None:
            # own stream; every update ends with a device synchronisation, so the next
            # pair never overwrites buffers the network still reads
            net = U.UNet(nc, nc, nc)
            net.set_params(params)
        for _ in range(args.steps_per_scene):
            if pair >= args.pairs or time.perf_counter() - t0 >= args.max_seconds:
                break
            generate = args.fixed_pairs <= 0 or pair < args.fixed_pairs
            if generate:
                ctx.step(1)
                rel, its = ctx.training_pair(inp, tgt, coarsening=args.coarsening, rel_tol=1e-6, max_iterations=20000)
                cg_its_total += its
            if scale is None:
                t_host = hip.to_host(tgt.value, cells)
                rms = float(np.sqrt(np.mean(t_host.astype(np.float64) ** 2)))
                scale = 2.0 ** math.ceil(math.log2(max(rms, 1e-6)))
                record["target_scale"] = scale
                print(f"target rms {rms:.3e}, training scale {scale}", flush=True)
            if generate:
                # store tgt / scale in the replay ring; each pair is weighted by 1 / mean(t^2)
                # (its loss is the relative squared error: targets span orders of magnitude)
                t_host = hip.to_host(tgt.value, cells) / np.float32(scale)
                ms = float(np.mean(t_host.astype(np.float64) ** 2))
                if not ms > 0.0:
                    continue
                slot = pair % R
                hip.copy_device(rep_inp.value + slot * cells * 8, inp.value, cells * 8)
                hip.to_device(rep_tgt.value + slot * cells * 4, t_host.astype(np.float32))
                rep_w[slot] = 1.0 / ms
                filled = min(filled + 1, R)
            frac = pair / max(1, args.pairs - 1)
            lr = args.lr_final + 0.5 * (args.lr - args.lr_final) * (1.0 + math.cos(math.pi * frac))
            for u in range(max(1, args.updates_per_pair)):
                slot = (pair % R) if u == 0 and generate else int(rng.integers(0, filled))
                hip.to_device(loss_d.value, np.zeros(1, np.float32))
                net.backward(rep_inp.value + slot * cells * 8, rep_tgt.value + slot * cells * 4, loss_d, weight=float(rep_w[slot]), centered=True)
                net.adam_step(lr)
                hip.synchronize()
                loss = float(hip.to_host(loss_d.value, 1)[0])
                ema = loss if ema is None else 0.98 * ema + 0.02 * loss
                updates += 1
            pair += 1
            if pair % 50 == 0:
                record["losses"].append([pair, ema])
            now = time.perf_counter()
            if now - last_print > 20.0:
                last_print = now
                print(f"pair {pair} updates {updates} scene {scene_index} loss(ema) {ema:.4e} lr {lr:.2e} cg its/pair {cg_its_total / pair:.0f} "
                      f"rel {rel:.1e} elapsed {now - t0:.0f}s", flush=True)
                if not math.isfinite(ema):
                    break
    out = net.params()
    net.destroy()
    ctx.destroy()
    s = last_layer_slice(count)
    out[s] *= np.float32(scale)
    out.astype("<f4").tofile(args.out)
    record.update({"pairs": pair, "updates": updates, "replay": R, "scenes": scene_index, "seconds": time.perf_counter() - t0, "final_loss_ema": ema,
                   "cg_iterations_per_pair": cg_its_total / max(pair, 1), "param_count": count,
                   "data": "random plumes at 128^3 (free-slip box), pre-projection divergence, FP64 CG to 1e-6"})
    json.dump(record, open(args.out + ".json", "w"), indent=1)
    print(json.dumps({k: v for k, v in record.items() if k != "losses"}), flush=True)


if __name__ == "__main__":
    main()
