"""Pins the oracle's smoothers (reference backend-cuda.cu:354-375, :398-407 and the
iteration loops :628-662) against an independent matrix formulation.

The 6-neighbour sum with ghost cells is written as S(x) = N x + s * x + k, where N is the
sparse adjacency of in-grid (or wrapped) neighbours, s counts copy ghosts minus reflected
ghosts per cell and k collects the constants (fixed values, outflow pressure, 2 * wall
velocity). One red-black iteration relaxes the red cells from x, then the black cells
from the half-updated vector; Jacobi relaxes every cell from x. Converged results are
checked against dense solves of the fixed-point system and against DCT/FFT solves of
the discrete Poisson problem.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.fft
import scipy.sparse as sp

from conftest import sf

abi = sf.abi
RBGS, JACOBI = 0, 1


def P(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@pytest.fixture(scope="module")
def lib(oracle):
    return oracle.load(abi)


def operator(shape, faces, periodic):
    """faces[k] = (mode, value) for k = 2*axis + side, mode in {'copy','reflect','const'};
    periodic[a] for axis a (0 = x). Returns N (sparse), s (vector), k (vector)."""
    nz, ny, nx = shape
    n = nx * ny * nz
    idx = np.arange(n).reshape(shape)
    coords = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")  # z, y, x
    ext = (nx, ny, nz)
    rows, cols = [], []
    s = np.zeros(n)
    k = np.zeros(n)
    for axis in range(3):
        npax = 2 - axis
        c = coords[npax]
        for side, d in ((0, -1), (1, 1)):
            q = c + d
            inside = (q >= 0) & (q < ext[axis])
            if periodic[axis]:
                q = q % ext[axis]
                inside = np.ones_like(inside)
            nb = [coords[0], coords[1], coords[2]]
            nb[npax] = np.where(inside, q, c)
            nbi = idx[nb[0], nb[1], nb[2]]
            rows.append(idx[inside])
            cols.append(nbi[inside])
            mode, value = faces[2 * axis + side]
            ghost = ~inside
            if mode == "copy":
                s[idx[ghost]] += 1
            elif mode == "reflect":
                s[idx[ghost]] -= 1
                k[idx[ghost]] += value
            else:
                k[idx[ghost]] += value
    r, cc = np.concatenate(rows), np.concatenate(cols)
    N = sp.csr_matrix((np.ones_like(r, dtype=np.float64), (r, cc)), shape=(n, n))
    return N, s, k


def matrix_iterate(shape, N, s, k, b, x, iterations, relax, smoother):
    nz, ny, nx = shape
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    red = ((xx + yy + zz) % 2 == 0).ravel()
    x = x.ravel().copy()
    b = b.ravel()
    for _ in range(iterations):
        if smoother == JACOBI:
            x = relax(b, N @ x + s * x + k)
            continue
        half = x.copy()
        half[red] = relax(b, N @ x + s * x + k)[red]
        new = half.copy()
        new[~red] = relax(b, N @ half + s * half + k)[~red]
        x = new
    return x.reshape(shape)


def scalar_faces(types, values):
    return [("const", float(np.float32(values[i]))) if t == abi.SCALAR_FIXED else ("copy", 0.0) for i, t in enumerate(types)]


def velocity_faces(types, wall_v, comp):
    out = []
    for i, t in enumerate(types):
        axis = i // 2
        if t == abi.FLOW_OUTFLOW or (t == abi.FLOW_FREE_SLIP and comp != axis):
            out.append(("copy", 0.0))
        else:
            out.append(("reflect", 2.0 * float(np.float32(wall_v[i][comp]))))  # the ABI stores float
    return out


def pressure_faces(types, press):
    return [("const", float(np.float32(press[i]))) if t == abi.FLOW_OUTFLOW else ("copy", 0.0) for i, t in enumerate(types)]


SCALAR_CASES = [
    ((5, 4, 6), [abi.SCALAR_ZERO_FLUX] * 6, [0] * 6),
    ((4, 6, 5), [abi.SCALAR_FIXED, abi.SCALAR_ZERO_FLUX, abi.SCALAR_PERIODIC, abi.SCALAR_PERIODIC, abi.SCALAR_FIXED, abi.SCALAR_FIXED], [1.5, 0, 0, 0, -2.0, 0.25]),
    ((3, 5, 7), [abi.SCALAR_PERIODIC] * 2 + [abi.SCALAR_ZERO_FLUX] * 4, [0] * 6),  # odd periodic x
]


@pytest.mark.parametrize("smoother", [RBGS, JACOBI])
@pytest.mark.parametrize("case", range(len(SCALAR_CASES)))
def test_scalar_smoother_matches_matrix_iteration(lib, case, smoother):
    shape, types, values = SCALAR_CASES[case]
    nz, ny, nx = shape
    periodic = [types[2 * a] == abi.SCALAR_PERIODIC for a in range(3)]
    N, s, k = operator(shape, scalar_faces(types, values), periodic)
    rng = np.random.default_rng(case)
    b = rng.random(shape)
    x0 = rng.random(shape)
    alpha = 0.37
    for iters in (1, 2, 7):
        x = x0.copy()
        lib.sfo_smooth_scalar(P(x), P(b), alpha, iters, smoother, nx, ny, nz, C.byref(abi.scalar_boundary(types, values)))
        ref = matrix_iterate(shape, N, s, k, b, x0, iters, lambda bb, S: (bb + alpha * S) / (1 + 6 * alpha), smoother)
        np.testing.assert_allclose(x, ref, rtol=1e-13, atol=1e-13)


FLOW_TYPES = [abi.FLOW_NO_SLIP, abi.FLOW_FREE_SLIP, abi.FLOW_FREE_SLIP, abi.FLOW_OUTFLOW, abi.FLOW_PERIODIC, abi.FLOW_PERIODIC]
WALL_V = [(0.5, -0.25, 0.125), (0.0, 0.0, 0.0), (0.3, 0.6, -0.9), (0, 0, 0), (0, 0, 0), (0, 0, 0)]
PRESS = [0.0, 0.0, 0.0, 0.75, 0.0, 0.0]


@pytest.mark.parametrize("comp", [0, 1, 2])
@pytest.mark.parametrize("smoother", [RBGS, JACOBI])
def test_velocity_smoother_matches_matrix_iteration(lib, comp, smoother):
    shape = (4, 5, 6)
    nz, ny, nx = shape
    fbd = abi.flow_boundary(FLOW_TYPES, velocity=WALL_V, pressure=PRESS)
    N, s, k = operator(shape, velocity_faces(FLOW_TYPES, WALL_V, comp), [False, False, True])
    rng = np.random.default_rng(10 + comp)
    b, x0 = rng.standard_normal(shape), rng.standard_normal(shape)
    alpha = 0.8
    x = x0.copy()
    lib.sfo_smooth_velocity(P(x), P(b), alpha, comp, 3, smoother, nx, ny, nz, C.byref(fbd))
    ref = matrix_iterate(shape, N, s, k, b, x0, 3, lambda bb, S: (bb + alpha * S) / (1 + 6 * alpha), smoother)
    np.testing.assert_allclose(x, ref, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("smoother", [RBGS, JACOBI])
def test_pressure_smoother_matches_matrix_iteration(lib, smoother):
    shape = (6, 5, 4)
    nz, ny, nx = shape
    fbd = abi.flow_boundary(FLOW_TYPES, velocity=WALL_V, pressure=PRESS)
    N, s, k = operator(shape, pressure_faces(FLOW_TYPES, PRESS), [False, False, True])
    rng = np.random.default_rng(7)
    div, p0 = rng.standard_normal(shape), np.zeros(shape)
    h2 = 0.01
    p = p0.copy()
    lib.sfo_smooth_pressure(P(p), P(div), h2, 4, smoother, nx, ny, nz, C.byref(fbd))
    ref = matrix_iterate(shape, N, s, k, div, p0, 4, lambda bb, S: (S - h2 * bb) / 6.0, smoother)
    np.testing.assert_allclose(p, ref, rtol=1e-13, atol=1e-13)


def test_converged_diffusion_equals_dense_solve(lib):
    shape, types, values = SCALAR_CASES[1]
    nz, ny, nx = shape
    periodic = [types[2 * a] == abi.SCALAR_PERIODIC for a in range(3)]
    N, s, k = operator(shape, scalar_faces(types, values), periodic)
    alpha = 0.9
    A = (1 + 6 * alpha) * np.eye(N.shape[0]) - alpha * (N.toarray() + np.diag(s))
    b = np.random.default_rng(2).random(shape)
    exact = np.linalg.solve(A, b.ravel() + alpha * k).reshape(shape)
    for smoother in (RBGS, JACOBI):
        x = b.copy()
        lib.sfo_smooth_scalar(P(x), P(b), alpha, 400, smoother, nx, ny, nz, C.byref(abi.scalar_boundary(types, values)))
        np.testing.assert_allclose(x, exact, rtol=0, atol=1e-10)


def test_converged_pressure_equals_dct_neumann_solve(lib):
    """All walls: the clamp ghosts make the 7-point Laplacian diagonal in DCT-II."""
    n, h = 12, 1.0 / 12
    shape = (n, n, n)
    rng = np.random.default_rng(5)
    div = rng.standard_normal(shape)
    div -= div.mean()  # compatibility of the Neumann problem
    lam1 = 2.0 * np.cos(np.pi * np.arange(n) / n) - 2.0
    lam = lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :]
    rhs = scipy.fft.dctn(h * h * div, type=2, norm="ortho")
    lam[0, 0, 0] = 1.0
    ph = rhs / lam
    ph[0, 0, 0] = 0.0
    exact = scipy.fft.idctn(ph, type=2, norm="ortho")
    fbd = abi.flow_boundary([abi.FLOW_FREE_SLIP] * 6)
    p = np.zeros(shape)
    lib.sfo_smooth_pressure(P(p), P(div), h * h, 1500, RBGS, n, n, n, C.byref(fbd))
    p -= p.mean()
    np.testing.assert_allclose(p, exact - exact.mean(), rtol=0, atol=1e-9 * np.abs(exact).max())


def test_converged_pressure_equals_fft_periodic_solve(lib):
    n, h = 8, 0.125
    shape = (n, n, n)
    div = np.random.default_rng(6).standard_normal(shape)
    div -= div.mean()
    lam1 = 2.0 * np.cos(2 * np.pi * np.arange(n) / n) - 2.0
    lam = lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :]
    lam[0, 0, 0] = 1.0
    ph = np.fft.fftn(h * h * div) / lam
    ph[0, 0, 0] = 0.0
    exact = np.real(np.fft.ifftn(ph))
    fbd = abi.flow_boundary([abi.FLOW_PERIODIC] * 6)
    p = np.zeros(shape)
    lib.sfo_smooth_pressure(P(p), P(div), h * h, 800, RBGS, n, n, n, C.byref(fbd))
    p -= p.mean()
    np.testing.assert_allclose(p, exact, rtol=0, atol=1e-9 * np.abs(exact).max())


def test_zero_divergence_gives_zero_pressure(lib):
    fbd = abi.flow_boundary([abi.FLOW_FREE_SLIP] * 6)
    p = np.zeros((4, 4, 4))
    lib.sfo_smooth_pressure(P(p), P(np.zeros((4, 4, 4))), 0.01, 10, RBGS, 4, 4, 4, C.byref(fbd))
    assert np.all(p == 0)


def test_one_red_black_iteration_by_hand(lib):
    """3 cells in x (1x1x3), zero-flux ends, alpha = 1: red cells 0 and 2, black cell 1.
    x0 = (0, 0, 0), b = (7, 14, 7). Red: x0' = (7 + 1*(0 + 0)) / 7 = 1 (ghost copies
    itself = 0, neighbour 0), x2' = 1. Black: x1' = (14 + (1 + 1 + 4 * 0)) / 7 = 16/7."""
    b = np.array([7.0, 14.0, 7.0]).reshape(1, 1, 3)
    x = np.zeros((1, 1, 3))
    sbd = abi.scalar_boundary([abi.SCALAR_ZERO_FLUX] * 6)
    lib.sfo_smooth_scalar(P(x), P(b), 1.0, 1, RBGS, 3, 1, 1, C.byref(sbd))
    np.testing.assert_allclose(x.ravel(), [1.0, 16.0 / 7.0, 1.0], rtol=1e-15)
