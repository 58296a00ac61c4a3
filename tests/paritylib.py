"""Shared helpers for the GPU-vs-oracle parity tests."""
import numpy as np

from conftest import sf, sf_oracle

abi = sf.abi
FLOW_KINDS = {"pressure": abi.VIEW_PRESSURE, "divergence": abi.VIEW_DIVERGENCE, "speed": abi.VIEW_VELOCITY_MAGNITUDE}


def gpu_fields(ctx, desc):
    out = {k: ctx.read(v).astype(np.float64) for k, v in FLOW_KINDS.items()}
    for name, arr in zip(("vx", "vy", "vz"), ctx.read(abi.VIEW_VELOCITY)):
        out[name] = arr.astype(np.float64)
    for s in range(desc.n_scalars):
        out[f"scalar{s + 1}"] = ctx.read(abi.VIEW_SCALAR_DATA, s + 1).astype(np.float64)
    return out


def oracle_fields(sim, desc):
    out = {k: sim.read(v) for k, v in FLOW_KINDS.items()}
    for name, arr in zip(("vx", "vy", "vz"), sim.read(abi.VIEW_VELOCITY)):
        out[name] = arr
    for s in range(desc.n_scalars):
        out[f"scalar{s + 1}"] = sim.read(abi.VIEW_SCALAR_DATA, s + 1)
    return out


def normwise(a, b):
    """max |a - b| / max |b| per field (0 when both are identically zero)."""
    res = {}
    for k in b:
        scale = np.max(np.abs(b[k]))
        diff = np.max(np.abs(a[k] - b[k]))
        res[k] = 0.0 if diff == 0.0 else float(diff / max(scale, 1e-300))
    return res


class Inputs:
    """Seeded host inputs for every field the ABI can set."""

    def __init__(self, desc, seed, velocity_scale=0.0, source_scale=1.0, force_scale=1.0, smooth=False):
        """White noise per cell, or (smooth=True) the same noise after a 5^3 box filter."""
        rng = np.random.default_rng(seed)
        shape = desc.shape
        f = (lambda a: sf.scenes.box_filter5(a) * 3.0) if smooth else (lambda a: a)
        self.sources = [f(rng.standard_normal(shape) * source_scale).astype(np.float32) for _ in range(desc.n_scalars)]
        self.densities = [f(rng.random(shape)).astype(np.float32) for _ in range(desc.n_scalars)]
        self.forces = [tuple(f(rng.standard_normal(shape) * force_scale).astype(np.float32) for _ in range(3)) for _ in range(desc.n_vectors)]
        self.velocity = tuple(f(rng.standard_normal(shape) * velocity_scale).astype(np.float32) for _ in range(3)) if velocity_scale else None

    def seed_gpu(self, ctx):
        for i, (s, d) in enumerate(zip(self.sources, self.densities)):
            ctx.update_scalar_field(i + 1, ctx.upload(d))
            ctx.update_scalar_field_source(i + 1, ctx.upload(s))
        for i, f in enumerate(self.forces):
            ctx.update_vector_field(i + 1, *(ctx.upload(c) for c in f))
        if self.velocity is not None:
            ctx.set_flow_velocity(*(ctx.upload(c) for c in self.velocity))

    def seed_oracle(self, sim):
        for i, (s, d) in enumerate(zip(self.sources, self.densities)):
            sim.update_scalar_field(i + 1, d)
            sim.update_scalar_field_source(i + 1, s)
        for i, f in enumerate(self.forces):
            sim.update_vector_field(i + 1, *f)
        if self.velocity is not None:
            sim.set_flow_velocity(*self.velocity)


def run_gpu(desc, seeder, steps, options=None):
    ctx = sf.Context(desc, options=options)
    try:
        seeder(ctx)
        ctx.step(steps)
        return gpu_fields(ctx, desc)
    finally:
        ctx.destroy()


def run_oracle(desc, seeder, steps, precision, smoother=sf_oracle.RBGS):
    sim = sf_oracle.OracleSim(abi, desc, precision=precision, smoother=smoother)
    try:
        seeder(sim)
        sim.step(steps)
        return oracle_fields(sim, desc)
    finally:
        sim.destroy()


def report(tag, errs):
    print(tag + " " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
