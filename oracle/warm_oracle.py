"""FP64 host oracle of the learned pressure warm start (test infrastructure only).

Restates stable-fluids_amd/csrc/sf_warm.hip in numpy: the pressure operator of a step
(reference backend-cuda.cu:374-384, ghosts by load_pressure :96-128: walls copy the
cell, outflow faces give their pressure, periodic axes wrap), the block-mean
restriction (f^3 cells) and wall mask of the network input, the trilinear prolongation
of its output, the guess residual test and the fallback rule. Arrays are [z][y][x].

Pinned by tests/test_warm_oracle.py: the residual equals 6 (x_jacobi - x) from one
Jacobi pass of the C oracle (oracle/sf_oracle.cpp), whose ghosts are pinned against
the reference rules; prolongation reproduces linear fields away from the clamped edge
cells and constants everywhere; restriction of a constant is that constant.
"""
import numpy as np

from . import unet_oracle as uo

FLOW_OUTFLOW = 2
FLOW_PERIODIC = 3
FACES = ("x_minus", "x_plus", "y_minus", "y_plus", "z_minus", "z_plus")


def pressure_rules(fb):
    """[(kind, value)] per face (kind 'copy' | 'const' | 'periodic'), faces x-, x+, y-, y+, z-, z+."""
    out = []
    for name in FACES:
        f = getattr(fb, name)
        if f.type == FLOW_PERIODIC:
            out.append(("periodic", 0.0))
        elif f.type == FLOW_OUTFLOW:
            out.append(("const", float(np.float32(f.pressure))))
        else:
            out.append(("copy", 0.0))
    return out


def neighbour_sum(p, rules, homogeneous=False):
    """Sum of the 6 ghosted neighbours of every cell."""
    s = np.zeros_like(p, dtype=np.float64)
    for f, (kind, value) in enumerate(rules):
        axis = 2 - (f >> 1)  # x is the last array axis
        plus = f & 1
        if kind == "periodic":
            s += np.roll(p, -1 if plus else 1, axis=axis)
            continue
        shifted = np.roll(p, -1 if plus else 1, axis=axis).astype(np.float64)
        edge = [slice(None)] * 3
        edge[axis] = -1 if plus else 0
        if kind == "copy":
            shifted[tuple(edge)] = p[tuple(edge)]
        else:
            shifted[tuple(edge)] = 0.0 if homogeneous else value
        s += shifted
    return s


def pressure_residual(p, div, h2, rules):
    """b - A p with b = -h2 div and (A p) = 6 p - (ghosted neighbour sum)."""
    p = np.asarray(p, dtype=np.float64)
    return -h2 * np.asarray(div, dtype=np.float64) + neighbour_sum(p, rules) - 6.0 * p


def apply_homogeneous(p, rules):
    return 6.0 * p - neighbour_sum(np.asarray(p, dtype=np.float64), rules, homogeneous=True)


def restrict(a, f=2):
    """Mean over blocks of f^3 cells."""
    z, y, x = a.shape
    return np.asarray(a, dtype=np.float64).reshape(z // f, f, y // f, f, x // f, f).mean(axis=(1, 3, 5))


def wall_mask(shape, rules):
    m = np.zeros(shape)
    for f, (kind, _) in enumerate(rules):
        if kind != "copy":
            continue
        axis = 2 - (f >> 1)
        edge = [slice(None)] * 3
        edge[axis] = -1 if (f & 1) else 0
        m[tuple(edge)] = 1.0
    return m


def prolong(q, f=2):
    """Cell-centred trilinear interpolation to f times the resolution, clamped at the edges."""
    out = np.asarray(q, dtype=np.float64)
    for axis in range(3):
        n = out.shape[axis]
        c = np.clip((np.arange(f * n) + 0.5) / f - 0.5, 0.0, n - 1)
        i0 = np.floor(c).astype(int)
        i1 = np.minimum(i0 + 1, n - 1)
        t = c - i0
        shape = [1, 1, 1]
        shape[axis] = f * n
        t = t.reshape(shape)
        out = np.take(out, i0, axis=axis) * (1 - t) + np.take(out, i1, axis=axis) * t
    return out


def guess(div, params, h, rules, rnd=uo.identity, f=2):
    """Warm-start guess p0 on the full grid (network on the grid coarsened f times) and
    its (sigma, network output q)."""
    dc = restrict(div, f)
    sigma = float(np.sqrt(np.mean(dc * dc)))
    inv = 1.0 / sigma if sigma > 0 else 0.0
    inp = np.stack([dc * inv, wall_mask(dc.shape, rules)], axis=-1)
    q, _ = uo.forward(params, inp, rnd)
    return (f * h) ** 2 * sigma * prolong(q, f), sigma, q


def guess_test(p0, div, h2, rules, ratio=10.0):
    """(||b - A p0||, ||b||, fallback) with fallback = guess residual > ratio * zero residual."""
    g = float(np.sqrt(np.sum(pressure_residual(p0, div, h2, rules) ** 2)))
    z = float(np.sqrt(np.sum(pressure_residual(np.zeros_like(p0), div, h2, rules) ** 2)))
    return g, z, not (g <= ratio * z)
