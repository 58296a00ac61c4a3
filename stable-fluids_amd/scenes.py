"""Scene builders for the workloads the library is measured and checked on.

All fields are float32 arrays shaped (nz, ny, nx) (x fastest, as the C ABI stores them).
Cell (i, j, k) is 0-based with centre ((i+.5)h, (j+.5)h, (k+.5)h); +y is up.

plume(n): a sphere of radius r cells at integer centre (cx, cy, cz), membership test
(i-cx)^2 + (j-cy)^2 + (k-cz)^2 <= r^2; density source `rate` and upward force `force`
inside it, zero elsewhere. Walls: free-slip on all six faces for the flow, zero-flux for
the density (Stam's set_bnd: normal velocity negated, density copied).

noise_velocity: i.i.d. N(0, 1) per component and cell from a counter-based generator
(splitmix64 of the counter, Box-Muller on two 53-bit uniforms), then a 5^3 box average
with edge replication, so the field starts divergent but smooth.
"""
import numpy as np

from . import abi

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(counter):
    """splitmix64 output for counter values (uint64 array): mix((counter + 1) * golden)."""
    with np.errstate(over="ignore"):
        z = (counter + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def normal_noise(count, seed, stream):
    """`count` standard normals for (seed, stream); deterministic on any host."""
    base = (np.uint64(seed) << np.uint64(40)) ^ (np.uint64(stream) << np.uint64(34))
    idx = np.arange(count, dtype=np.uint64) * np.uint64(2) + base
    u1 = (splitmix64(idx) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u2 = (splitmix64(idx + np.uint64(1)) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def box_filter5(a):
    """5^3 box average with edge replication (clamped at the walls)."""
    p = np.pad(a, 2, mode="edge")
    for axis in range(3):
        c = np.cumsum(p, axis=axis, dtype=np.float64)
        c = np.concatenate([np.zeros_like(np.take(c, [0], axis=axis)), c], axis=axis)
        n = p.shape[axis] - 4
        p = (np.take(c, np.arange(5, n + 5), axis=axis) - np.take(c, np.arange(0, n), axis=axis)) / 5.0
    return p


def noise_velocity(shape, seed=0, sigma=1.0):
    """Three components of box-filtered N(0, sigma^2) noise (streams 0, 1, 2 of `seed`),
    generated concurrently (numpy releases the interpreter lock in these loops)."""
    from concurrent.futures import ThreadPoolExecutor
    nz, ny, nx = shape

    def comp(c):
        return box_filter5(normal_noise(nx * ny * nz, seed, c).reshape(shape) * sigma).astype(np.float32)

    with ThreadPoolExecutor(max_workers=3) as pool:
        return list(pool.map(comp, range(3)))


def sphere_mask(shape, centre, radius):
    nz, ny, nx = shape
    cx, cy, cz = centre
    k, j, i = np.ogrid[0:nz, 0:ny, 0:nx]
    return (i - cx) ** 2 + (j - cy) ** 2 + (k - cz) ** 2 <= radius * radius


def plume_desc(n, diffuse_iterations, pressure_iterations, dt=0.1, viscosity=1e-4, diffusion=1e-4,
               cell_size=None, flow_types=None, scalar_types=None, dissipation=0.0, shape=None):
    nx, ny, nz = (n, n, n) if shape is None else shape
    h = 1.0 / max(nx, ny, nz) if cell_size is None else cell_size
    flow = abi.flow_boundary(flow_types or [abi.FLOW_FREE_SLIP] * 6)
    sb = abi.scalar_boundary(scalar_types or [abi.SCALAR_ZERO_FLUX] * 6)
    return abi.Desc(nx, ny, nz, h, dt, viscosity, diffuse_iterations, pressure_iterations, flow,
                    scalars=[abi.scalar_desc(diffusion, dissipation, 0.0, sb)], vectors=[abi.force_desc()])


def plume_fields(shape, centre, radius, rate=100.0, force=5.0):
    """(density source, force x, force y, force z) of the plume sphere."""
    m = sphere_mask(shape, centre, radius)
    src = np.where(m, np.float32(rate), np.float32(0)).astype(np.float32)
    fy = np.where(m, np.float32(force), np.float32(0)).astype(np.float32)
    zero = np.zeros(shape, np.float32)
    return src, zero, fy, zero.copy()


class Scene:
    """A context description plus the host fields that seed it."""

    def __init__(self, name, desc, source, force, velocity=None):
        self.name = name
        self.desc = desc
        self.source = source
        self.force = force
        self.velocity = velocity

    @property
    def shape(self):
        return self.desc.shape


def plume_scene(n=64, steps_scale=1, diffuse_iterations=20, pressure_iterations=40, noise_seed=None, force=5.0, **kw):
    """The plume at grid n^3: centre (n/2, 3n/32, n/2), radius n/16, rate 100, upward
    force `force` (5 by default).

    n=64 gives the 64^3 oracle workload (centre (32,6,32), radius 4); n=256 the 256^3
    throughput workload (centre (128,24,128), radius 16) when noise_seed=0 and 40/80
    iterations; n=512 the 512^3 precision workload (centre (256,48,256), radius 32).
    """
    desc = plume_desc(n, diffuse_iterations, pressure_iterations, **kw)
    centre = (n // 2, (3 * n) // 32, n // 2)
    src, fx, fy, fz = plume_fields(desc.shape, centre, n // 16, 100.0, force)
    vel = noise_velocity(desc.shape, noise_seed) if noise_seed is not None else None
    return Scene(f"plume{n}", desc, src, (fx, fy, fz), vel)


def reference_benchmark_scene(nx=128, ny=160, nz=128, diffuse_iterations=24, pressure_iterations=96, cell_size=0.01,
                              dt=1.0 / 90.0, viscosity=0.00012, diffusion=0.00005, dissipation=0.35):
    """The scene of the reference's benchmark driver (benchmark.cu:16-27, :82-155):
    periodic x/z, no-slip floor, outflow top, ellipsoidal plume with swirling force."""
    flow = abi.flow_boundary([abi.FLOW_PERIODIC, abi.FLOW_PERIODIC, abi.FLOW_NO_SLIP, abi.FLOW_OUTFLOW,
                              abi.FLOW_PERIODIC, abi.FLOW_PERIODIC])
    sb = abi.scalar_boundary([abi.SCALAR_PERIODIC, abi.SCALAR_PERIODIC, abi.SCALAR_ZERO_FLUX, abi.SCALAR_ZERO_FLUX,
                              abi.SCALAR_PERIODIC, abi.SCALAR_PERIODIC])
    desc = abi.Desc(nx, ny, nz, cell_size, dt, viscosity, diffuse_iterations, pressure_iterations, flow,
                    scalars=[abi.scalar_desc(diffusion, dissipation, 0.0, sb)], vectors=[abi.force_desc()])
    f32 = np.float32
    ex, ey, ez = f32(nx) * f32(cell_size), f32(ny) * f32(cell_size), f32(nz) * f32(cell_size)
    sx, sy, sz = ex * f32(0.5), ey * f32(0.16), ez * f32(0.5)
    rx, ry, rz = ex * f32(0.07), ey * f32(0.06), ez * f32(0.07)
    k, j, i = np.meshgrid(np.arange(nz, dtype=f32), np.arange(ny, dtype=f32), np.arange(nx, dtype=f32), indexing="ij")
    px, py, pz = (i + f32(0.5)) * f32(cell_size), (j + f32(0.5)) * f32(cell_size), (k + f32(0.5)) * f32(cell_size)
    dx, dy, dz = (px - sx) / rx, (py - sy) / ry, (pz - sz) / rz
    r2 = dx * dx + dy * dy + dz * dz
    inside = r2 <= f32(1.0)
    plume = np.exp(f32(-2.2) * r2).astype(f32)
    src = np.where(inside, f32(28.0) * plume, f32(0)).astype(f32)
    fx = np.where(inside, f32(1.8) * plume * np.sin((pz / (ez + f32(1e-6))) * f32(6.2831853)), f32(0)).astype(f32)
    fy = np.where(inside, f32(7.5) * plume, f32(0)).astype(f32)
    fz = np.where(inside, f32(1.4) * plume * np.cos((px / (ex + f32(1e-6))) * f32(6.2831853)), f32(0)).astype(f32)
    return Scene("reference-benchmark", desc, src, (fx, fy, fz))


def viewer_plume_scene(nx=96, ny=128, nz=96, diffuse_iterations=24, pressure_iterations=96, cell_size=0.01,
                       dt=1.0 / 90.0, viscosity=0.00012, diffusion=0.00005, dissipation=0.35):
    """The visualizer's plume scene (vulkan-app/scene_plume.ixx:34-50, scene_plume.cpp:185-289):
    periodic x/z, no-slip floor, outflow top; a source sphere of radius 6h with source 32 w
    and upward force 7.6 w (w = 1 - d^2/r^2), plus the masks of its animated forcing: the
    drift mask (radius 2.1 r, centred 1.35 r above the source) and the swirl masks (radius
    1.65 r, 0.9 r above; the unit tangent (-dz, dx)/|(dx, dz)| times the weight). Returns
    (scene, (drift, swirl_x, swirl_z)); the forces' x and z start at zero."""
    flow = abi.flow_boundary([abi.FLOW_PERIODIC, abi.FLOW_PERIODIC, abi.FLOW_NO_SLIP, abi.FLOW_OUTFLOW,
                              abi.FLOW_PERIODIC, abi.FLOW_PERIODIC])
    sb = abi.scalar_boundary([abi.SCALAR_PERIODIC, abi.SCALAR_PERIODIC, abi.SCALAR_ZERO_FLUX, abi.SCALAR_ZERO_FLUX,
                              abi.SCALAR_PERIODIC, abi.SCALAR_PERIODIC])
    desc = abi.Desc(nx, ny, nz, cell_size, dt, viscosity, diffuse_iterations, pressure_iterations, flow,
                    scalars=[abi.scalar_desc(diffusion, dissipation, 0.0, sb)], vectors=[abi.force_desc()])
    f32 = np.float32
    h = f32(cell_size)
    sx, sy, sz = f32(nx) * h * f32(0.5), f32(ny) * h * f32(0.13), f32(nz) * h * f32(0.5)
    r = h * f32(6.0)
    swirl_y, drift_y = sy + r * f32(0.9), sy + r * f32(1.35)
    k, j, i = np.meshgrid(np.arange(nz, dtype=f32), np.arange(ny, dtype=f32), np.arange(nx, dtype=f32), indexing="ij")
    px, py, pz = (i + f32(0.5)) * h, (j + f32(0.5)) * h, (k + f32(0.5)) * h

    def radial_weight(cy, radius):
        dx, dy, dz = px - sx, py - cy, pz - sz
        r2 = radius * radius
        d2 = dx * dx + dy * dy + dz * dz
        return np.where((d2 >= r2) | (r2 <= f32(0)), f32(0), f32(1) - d2 / r2).astype(f32)

    w_src = radial_weight(sy, r)
    w_swirl = radial_weight(swirl_y, r * f32(1.65))
    w_drift = radial_weight(drift_y, r * f32(2.1))
    dx, dz = px - sx, pz - sz
    radial = np.sqrt(dx * dx + dz * dz).astype(f32)
    with np.errstate(divide="ignore"):
        inv = np.where(radial > f32(1e-5), f32(1) / radial, f32(0)).astype(f32)
    swirl_x = (-dz * inv * w_swirl).astype(f32)
    swirl_z = (dx * inv * w_swirl).astype(f32)
    zero = np.zeros(desc.shape, f32)
    scene = Scene("viewer-plume", desc, (f32(32.0) * w_src).astype(f32), (zero, (f32(7.6) * w_src).astype(f32), zero.copy()))
    return scene, (w_drift, swirl_x, swirl_z)


BATCH_SIZES = [(32, 32, 32), (48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 128, 128), (128, 64, 64)]


def batch_plume_scenes(count=64, seed=0, diffuse_iterations=40, pressure_iterations=80):
    """`count` plume simulations of uneven sizes: (nx, ny, nz) drawn from BATCH_SIZES, a
    random centre, radius U[2, 8] cells, upward force U[1, 10]; free-slip box, h = 1/max(n),
    dt = 0.1, nu = kappa = 1e-4."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nx, ny, nz = BATCH_SIZES[int(rng.integers(0, len(BATCH_SIZES)))]
        desc = plume_desc(max(nx, ny, nz), diffuse_iterations, pressure_iterations, shape=(nx, ny, nz))
        radius = float(rng.uniform(2.0, 8.0))
        centre = (int(rng.integers(int(radius) + 1, nx - int(radius) - 1)), int(rng.integers(int(radius) + 1, max(int(radius) + 2, ny // 4))),
                  int(rng.integers(int(radius) + 1, nz - int(radius) - 1)))
        src, fx, fy, fz = plume_fields(desc.shape, centre, radius, 100.0, float(rng.uniform(1.0, 10.0)))
        out.append(Scene(f"batch{i}", desc, src, (fx, fy, fz), None))
    return out


def animated_force(masks, step, animation):
    """Host restatement of the per-step forcing (vulkan-app/scene_plume.cpp:327-334), FP32:
    force x and z of step `step` from the masks and an abi.ForceAnimation's constants."""
    f32 = np.float32
    drift, swirl_x, swirl_z = masks
    phase = f32(step) * f32(animation.phase_rate)
    ax = f32(animation.drift_x_amplitude) * f32(np.sin(phase))
    az = f32(animation.drift_z_amplitude) * f32(np.cos(phase * f32(animation.drift_z_rate)))
    sw = f32(animation.swirl_amplitude) * f32(np.cos(phase * f32(animation.swirl_rate)))
    return (ax * drift + sw * swirl_x).astype(f32), (az * drift + sw * swirl_z).astype(f32)


def seed_context(ctx, scene):
    """Uploads the scene's source, force (and initial velocity) into a Context."""
    ctx.update_scalar_field_source(ctx.scalar_handles[0], ctx.upload(scene.source))
    fx, fy, fz = scene.force
    ctx.update_vector_field(ctx.vector_handles[0], ctx.upload(fx), ctx.upload(fy), ctx.upload(fz))
    if scene.velocity is not None:
        vx, vy, vz = scene.velocity
        ctx.set_flow_velocity(ctx.upload(vx), ctx.upload(vy), ctx.upload(vz))


def seed_oracle(sim, scene):
    sim.update_scalar_field_source(1, scene.source)
    sim.update_vector_field(1, *scene.force)
    if scene.velocity is not None:
        sim.set_flow_velocity(*scene.velocity)
