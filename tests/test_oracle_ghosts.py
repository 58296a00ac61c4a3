"""Pins the oracle's ghost-value rules (reference backend-cuda.cu:62-189) with hand-computed
tables. Field on a 3x3x3 grid: f[z, y, x] = 100 z + 10 y + x + 1, so every value names
its cell. Each expected number below is worked out by hand from the rule text, not by
re-running the rule."""
import numpy as np
import pytest

from conftest import sf

abi = sf.abi
N = 3


@pytest.fixture(scope="module")
def lib(oracle):
    return oracle.load(abi)


@pytest.fixture(scope="module")
def field():
    k, j, i = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return np.ascontiguousarray((100 * k + 10 * j + i + 1).astype(np.float64))


def ptr(a):
    import ctypes as C
    return a.ctypes.data_as(C.POINTER(C.c_double))


def sb(types, values=None):
    return abi.scalar_boundary(types, values)


Z, FX, PER = abi.SCALAR_ZERO_FLUX, abi.SCALAR_FIXED, abi.SCALAR_PERIODIC


@pytest.mark.parametrize("types,values,point,expect", [
    ([Z] * 6, None, (1, 1, 1), 112.0),                 # interior
    ([Z] * 6, None, (-1, 1, 1), 111.0),                # zero flux: clamp x to 0
    ([Z] * 6, None, (3, 2, 1), 123.0),                 # clamp x to 2
    ([Z] * 6, None, (1, -5, 2), 202.0),                # clamp far outside
    ([FX, Z, Z, Z, Z, Z], [7.5, 0, 0, 0, 0, 0], (-1, 1, 1), 7.5),
    ([FX, Z, Z, Z, Z, Z], [7.5, 0, 0, 0, 0, 0], (3, 1, 1), 113.0),   # other face zero flux
    ([PER, PER, Z, Z, Z, Z], None, (-1, 1, 1), 113.0),               # wrap to x=2
    ([PER, PER, Z, Z, Z, Z], None, (3, 1, 1), 111.0),
    ([PER, PER, Z, Z, Z, Z], None, (-4, 1, 1), 113.0),               # -4 mod 3 = 2
    ([PER, PER, Z, Z, Z, Z], None, (7, 0, 0), 2.0),                  # 7 mod 3 = 1
    # corner precedence: x is resolved first
    ([FX, Z, FX, Z, Z, Z], [7.5, 0, 2.5, 0, 0, 0], (-1, -1, 1), 7.5),
    ([Z, Z, FX, Z, Z, Z], [0, 0, 2.5, 0, 0, 0], (-1, -1, 1), 2.5),  # x clamps, y fixed
    ([Z, Z, Z, Z, Z, FX], [0, 0, 0, 0, 0, 4.0], (-1, -1, 3), 4.0),
    ([PER, PER, PER, PER, Z, Z], None, (-1, 3, -1), 3.0),            # (2, 0, clamp 0) -> 0+0+2+1
])
def test_scalar_ghosts(lib, field, types, values, point, expect):
    b = sb(types, values)
    import ctypes as C
    got = lib.sfo_load_scalar(ptr(field), *point, N, N, N, C.byref(b))
    assert got == expect


NS, FS, OUT, FP = abi.FLOW_NO_SLIP, abi.FLOW_FREE_SLIP, abi.FLOW_OUTFLOW, abi.FLOW_PERIODIC
WALL_V = [(0.5, 0.25, 0.125), (1.0, 2.0, 3.0), (-1.0, 0.5, 0.0), (0.0, 0.0, 0.75), (0.25, 0.0, -0.5), (0.0, 1.5, 0.0)]
PRESS = [0.0, 0.0, 0.0, 3.0, 0.0, -2.0]


def fb(types):
    return abi.flow_boundary(types, velocity=WALL_V, pressure=PRESS)


@pytest.mark.parametrize("types,point,expect", [
    ([FS] * 6, (1, 1, 1), 112.0),
    ([FS] * 6, (-1, 1, 1), 111.0),                    # walls clamp (Neumann)
    ([NS] * 6, (1, 1, 3), 212.0),
    ([FS, FS, FS, OUT, FS, FS], (1, 3, 1), 3.0),      # outflow: face pressure
    ([FS, FS, FS, OUT, FS, OUT], (1, 1, 3), -2.0),
    ([FP, FP, FS, FS, FS, FS], (-1, 1, 1), 113.0),
    # corner: x clamps then y outflow returns the face pressure
    ([FS, FS, FS, OUT, FS, FS], (-1, 3, 1), 3.0),
    # corner: y outflow first would not matter, x is first and clamps, z clamps: value
    ([FS] * 6, (-1, 3, -1), 21.0),                    # (0, 2, 0) -> 0 + 20 + 0 + 1
])
def test_pressure_ghosts(lib, field, types, point, expect):
    import ctypes as C
    b = fb(types)
    assert lib.sfo_load_pressure(ptr(field), *point, N, N, N, C.byref(b)) == expect


@pytest.mark.parametrize("types,comp,point,expect", [
    # free slip x-: normal component reflects about the wall velocity, tangential copies
    ([FS] * 6, 0, (-1, 1, 1), 2 * 0.5 - 111.0),
    ([FS] * 6, 1, (-1, 1, 1), 111.0),
    ([FS] * 6, 2, (3, 1, 1), 113.0),
    ([FS] * 6, 0, (3, 1, 1), 2 * 1.0 - 113.0),
    # no slip: every component reflects
    ([NS] * 6, 1, (-1, 1, 1), 2 * 0.25 - 111.0),
    ([NS] * 6, 2, (1, -1, 2), 2 * 0.0 - 202.0),
    ([NS] * 6, 0, (1, 3, 2), 2 * 0.0 - 222.0),
    # outflow copies the interior for every component
    ([FS, FS, FS, OUT, FS, FS], 1, (1, 3, 1), 122.0),
    # periodic wraps
    ([FP, FP, FS, FS, FS, FS], 0, (-1, 1, 1), 113.0),
    # corner: x wall resolved first, interior at (0, clamp y, clamp z)
    ([NS, NS, FS, FS, FS, FS], 1, (-1, -1, 1), 2 * 0.25 - 101.0),
    # x periodic wraps to 2, then y wall: interior (2, 0, 1), free-slip tangential copies
    ([FP, FP, FS, FS, FS, FS], 0, (-1, -1, 1), 103.0),
    # later axes are clamped, not wrapped: x wall, y periodic -> interior (0, 0, 1), not (0, 2, 1)
    ([FS, FS, FP, FP, FS, FS], 1, (-1, -1, 1), 101.0),
    ([NS, NS, FP, FP, FS, FS], 0, (-1, -1, 1), 2 * 0.5 - 101.0),
])
def test_velocity_ghosts(lib, field, types, comp, point, expect):
    import ctypes as C
    b = fb(types)
    assert lib.sfo_load_velocity(ptr(field), comp, *point, N, N, N, C.byref(b)) == expect
