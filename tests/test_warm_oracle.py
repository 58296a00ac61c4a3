"""The warm-start oracle (oracle/warm_oracle.py) pinned on the CPU.

* the pressure residual b - A p equals 6 (x_jacobi - x) from one Jacobi pass of the C
  oracle (whose ghost rules follow the reference's load_pressure) on mixed faces:
  periodic, walls and an outflow face with nonzero pressure;
* prolongation reproduces constants everywhere and linear fields away from the
  clamped edge cells; restriction of a constant is that constant;
* the fallback rule: zero guess never falls back, a guess 11x worse than zero does,
  a NaN guess does.
"""
import ctypes as C

import numpy as np

from conftest import sf
from oracle import sf_oracle, warm_oracle as wo

abi = sf.abi


def mixed_flow():
    return abi.flow_boundary([abi.FLOW_PERIODIC, abi.FLOW_PERIODIC, abi.FLOW_NO_SLIP, abi.FLOW_OUTFLOW,
                              abi.FLOW_FREE_SLIP, abi.FLOW_NO_SLIP], pressure=[0, 0, 0, 0.5, 0, 0])


def test_residual_matches_c_oracle_jacobi_pass():
    lib = sf_oracle.load(abi)
    fb = mixed_flow()
    shape = (5, 6, 8)
    rng = np.random.default_rng(1)
    p = rng.standard_normal(shape)
    div = rng.standard_normal(shape)
    h2 = 1.0 / 64
    x = p.copy()
    nz, ny, nx = shape
    lib.sfo_smooth_pressure(sf_oracle.dptr(x), sf_oracle.dptr(div), h2, 1, sf_oracle.JACOBI, nx, ny, nz, C.byref(fb))
    ref = 6.0 * (x - p)
    got = wo.pressure_residual(p, div, h2, wo.pressure_rules(fb))
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_prolong_and_restrict():
    q = np.full((3, 4, 5), 2.5)
    assert np.allclose(wo.prolong(q), 2.5, rtol=0, atol=1e-15)
    ic = np.arange(6, dtype=np.float64)
    q = np.broadcast_to(ic[None, None, :] * 3.0, (2, 2, 6))
    p = wo.prolong(q)
    xf = np.arange(12)
    expect = 3.0 * (0.5 * xf - 0.25)
    assert np.allclose(p[0, 0, 1:-1], expect[1:-1], atol=1e-12)
    assert p[0, 0, 0] == 0.0 and p[0, 0, -1] == 15.0  # clamped edge cells
    assert np.allclose(wo.restrict(np.full((4, 6, 8), -1.5)), -1.5)
    assert np.allclose(wo.restrict(np.full((8, 8, 12), 0.75), 4), 0.75)
    q = np.broadcast_to(ic[None, None, :] * 2.0, (2, 2, 6))
    p4 = wo.prolong(q, 4)
    xf = np.arange(24)
    expect4 = 2.0 * ((xf + 0.5) / 4 - 0.5)
    inner = (xf >= 2) & (xf <= 21)
    assert np.allclose(p4[0, 0, inner], expect4[inner], atol=1e-12)


def test_wall_mask_marks_wall_faces_only():
    rules = wo.pressure_rules(mixed_flow())
    m = wo.wall_mask((4, 4, 4), rules)
    assert m[:, 0, :].all() and m[0].all() and m[-1].all()   # no-slip floor, z walls
    assert not m[1:-1, -1, 1:-1].any()                        # outflow top is no wall
    assert not m[1:-1, 1:-1, 0].any()                         # periodic x


def test_fallback_rule():
    rules = wo.pressure_rules(mixed_flow())
    rng = np.random.default_rng(2)
    div = rng.standard_normal((4, 4, 4))
    zero = np.zeros((4, 4, 4))
    g, z, fb = wo.guess_test(zero, div, 1.0, rules)
    assert g == z and not fb
    # p0 with A p0 = -11 b (p0 = -11 x*, x* the solution) has residual 12 ||b|| > 10 ||b||
    big = np.full((4, 4, 4), 1e6)
    assert wo.guess_test(big, div, 1.0, rules)[2]
    nan = zero.copy()
    nan[0, 0, 0] = np.nan
    assert wo.guess_test(nan, div, 1.0, rules)[2]
