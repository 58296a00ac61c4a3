"""Pins the oracle's divergence, projection and whole-step composition
(reference backend-cuda.cu:386-434 and the step order :612-671)."""
import ctypes as C

import numpy as np
import pytest

from conftest import sf

abi = sf.abi
scenes = sf.scenes


def P(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.fixture(scope="module")
def lib(oracle):
    return oracle.load(abi)


def grid(shape):
    nz, ny, nx = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return i, j, k


def test_divergence_exact_on_affine_velocity_interior(lib):
    shape, h = (6, 7, 8), 0.2
    i, j, k = grid(shape)
    B = np.array([[0.5, -1.0, 2.0], [0.25, 1.5, -0.5], [1.0, 0.0, -2.5]])  # dv_c / dg_a
    v = [np.ascontiguousarray(0.1 * c + B[c, 0] * i + B[c, 1] * j + B[c, 2] * k, dtype=np.float64) for c in range(3)]
    div = np.zeros(shape)
    fbd = abi.flow_boundary([abi.FLOW_NO_SLIP] * 6)
    lib.sfo_divergence(P(div), P(v[0]), P(v[1]), P(v[2]), 8, 7, 6, h, C.byref(fbd))
    expect = (B[0, 0] + B[1, 1] + B[2, 2]) / h
    np.testing.assert_allclose(div[1:-1, 1:-1, 1:-1], expect, rtol=1e-12)


def test_divergence_telescopes_to_zero_with_free_slip_walls(lib):
    """Normal ghost = -interior at a still free-slip wall, so sum(div) telescopes to 0."""
    shape = (5, 6, 7)
    rng = np.random.default_rng(0)
    v = [rng.standard_normal(shape) for _ in range(3)]
    div = np.zeros(shape)
    fbd = abi.flow_boundary([abi.FLOW_FREE_SLIP] * 6)
    lib.sfo_divergence(P(div), P(v[0]), P(v[1]), P(v[2]), 7, 6, 5, 0.1, C.byref(fbd))
    assert abs(div.sum()) < 1e-11 * np.abs(div).sum()
    # a moving wall adds exactly its normal flux: x- wall velocity u gives +2u/(2h) per
    # boundary cell on that face: sum = ny*nz*u/h (sign from the minus side)
    fbd2 = abi.flow_boundary([abi.FLOW_FREE_SLIP] * 6, velocity=[(0.5, 0, 0)] + [(0, 0, 0)] * 5)
    lib.sfo_divergence(P(div), P(v[0]), P(v[1]), P(v[2]), 7, 6, 5, 0.1, C.byref(fbd2))
    assert div.sum() == pytest.approx(-6 * 5 * 0.5 / 0.1, rel=1e-10)


def test_project_subtracts_exact_gradient_of_affine_pressure(lib):
    shape, h = (6, 6, 6), 0.125
    i, j, k = grid(shape)
    p = np.ascontiguousarray(1.0 + 0.5 * i - 2.0 * j + 0.75 * k, dtype=np.float64)
    rng = np.random.default_rng(1)
    v = [rng.standard_normal(shape) for _ in range(3)]
    v0 = [a.copy() for a in v]
    fbd = abi.flow_boundary([abi.FLOW_FREE_SLIP] * 6)
    lib.sfo_project(P(v[0]), P(v[1]), P(v[2]), P(p), 6, 6, 6, h, C.byref(fbd))
    for c, g in enumerate((0.5, -2.0, 0.75)):
        np.testing.assert_allclose(v[c][1:-1, 1:-1, 1:-1], v0[c][1:-1, 1:-1, 1:-1] - g / h, rtol=1e-12, atol=1e-12)


def test_project_constant_pressure_with_walls_is_identity(lib):
    shape = (4, 5, 6)
    p = np.full(shape, 3.0)
    v = [np.random.default_rng(c).standard_normal(shape) for c in range(3)]
    v0 = [a.copy() for a in v]
    fbd = abi.flow_boundary([abi.FLOW_NO_SLIP] * 6)
    lib.sfo_project(P(v[0]), P(v[1]), P(v[2]), P(p), 6, 5, 4, 0.1, C.byref(fbd))
    for a, b in zip(v, v0):
        np.testing.assert_array_equal(a, b)


def step_desc(n=8, viscosity=0.0, diffusion=0.0, dissipation=0.0, d=3, p=5):
    flow = abi.flow_boundary([abi.FLOW_FREE_SLIP] * 6)
    return abi.Desc(n, n, n, 1.0 / n, 0.1, viscosity, d, p, flow, scalars=[abi.scalar_desc(diffusion, dissipation, 0.0)],
                    vectors=[abi.force_desc()])


def test_zero_viscosity_drops_advection_and_projects_the_forced_field(lib, oracle):
    """alpha <= 0: the advected velocity is never copied back (reference :621-627), so one
    step from rest is exactly: v = dt f; div; P pressure sweeps from 0; project. The
    density likewise keeps s + dt src (reference :656-658)."""
    desc = step_desc()
    n = 8
    rng = np.random.default_rng(2)
    f = [rng.standard_normal((n, n, n)).astype(np.float32) for _ in range(3)]
    src = rng.random((n, n, n)).astype(np.float32)
    sim = oracle.OracleSim(abi, desc, precision=64)
    sim.update_vector_field(1, *f)
    sim.update_scalar_field_source(1, src)
    sim.step(1)
    vx, vy, vz = sim.read(abi.VIEW_FLOW_VELOCITY)
    fbd = desc.desc.config.flow_boundary
    h, dt = np.float64(np.float32(1.0 / n)), np.float64(np.float32(0.1))
    v = [np.ascontiguousarray(dt * a.astype(np.float64)) for a in f]
    div = np.zeros((n, n, n))
    lib.sfo_divergence(P(div), P(v[0]), P(v[1]), P(v[2]), n, n, n, h, C.byref(fbd))
    p = np.zeros((n, n, n))
    lib.sfo_smooth_pressure(P(p), P(div), h * h, 5, 0, n, n, n, C.byref(fbd))
    lib.sfo_project(P(v[0]), P(v[1]), P(v[2]), P(p), n, n, n, h, C.byref(fbd))
    for got, want in zip((vx, vy, vz), v):
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(sim.read(abi.VIEW_FLOW_PRESSURE), p, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(sim.read(abi.VIEW_SCALAR_FIELD_DATA, 1), dt * src.astype(np.float64), rtol=1e-15)
    sim.destroy()


def test_dissipation_closed_form(oracle):
    """At rest, no source, kappa = 0: s_n = s_0 / (1 + dt k)^n (reference :664-667)."""
    desc = step_desc(dissipation=0.35)
    s0 = np.random.default_rng(3).random((8, 8, 8)).astype(np.float32)
    sim = oracle.OracleSim(abi, desc, precision=64)
    sim.update_scalar_field(1, s0)
    sim.step(4)
    factor = 1.0 / (1.0 + np.float64(np.float32(0.1)) * np.float64(np.float32(0.35)))
    np.testing.assert_allclose(sim.read(abi.VIEW_SCALAR_FIELD_DATA, 1), s0.astype(np.float64) * factor ** 4, rtol=1e-14)
    sim.destroy()


def test_rest_state_stays_at_rest(oracle):
    desc = step_desc(viscosity=1e-3, diffusion=1e-3)
    sim = oracle.OracleSim(abi, desc, precision=64)
    sim.step(3)
    for kind in (abi.VIEW_FLOW_PRESSURE, abi.VIEW_FLOW_DIVERGENCE, abi.VIEW_FLOW_VELOCITY_MAGNITUDE):
        assert np.all(sim.read(kind) == 0)
    sim.destroy()


def test_fp32_and_fp64_oracles_agree_on_short_horizon(oracle):
    scene = scenes.plume_scene(32, diffuse_iterations=10, pressure_iterations=20)
    out = {}
    for prec in (32, 64):
        sim = oracle.OracleSim(abi, scene.desc, precision=prec)
        scenes.seed_oracle(sim, scene)
        sim.step(2)
        out[prec] = [sim.read(abi.VIEW_SCALAR_FIELD_DATA, 1), sim.read(abi.VIEW_FLOW_PRESSURE)] + list(sim.read(abi.VIEW_FLOW_VELOCITY))
        sim.destroy()
    for a, b in zip(out[32], out[64]):
        assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(b))


def test_jacobi_and_rbgs_converge_to_the_same_projection(oracle):
    """nu = 0 (no diffusion) and many pressure iterations: both smoothers reach the same
    discrete pressure (up to the Neumann constant)."""
    res = []
    for smoother in (0, 1):
        desc = step_desc(p=3000)
        sim = oracle.OracleSim(abi, desc, precision=64, smoother=smoother)
        sim.update_vector_field(1, *(np.random.default_rng(c).standard_normal((8, 8, 8)).astype(np.float32) for c in range(3)))
        sim.step(1)
        p = sim.read(abi.VIEW_FLOW_PRESSURE)
        res.append(p - p.mean())
        sim.destroy()
    np.testing.assert_allclose(res[0], res[1], rtol=0, atol=1e-8 * np.abs(res[0]).max())
