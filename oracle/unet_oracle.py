"""FP64 host oracle of the pressure warm-start network (test infrastructure only).

Restates, in numpy float64, what the HIP network computes (stable-fluids_amd/csrc/
sf_unet.hip, sf_unet_kernels.hip; include/stable-fluids-unet.h): 3x3x3 convolutions with
zero padding, ReLU, 2x2x2 average pooling, nearest 2x upsampling, additive skips, MSE
loss, and the hand-derived backward pass. No counterpart in the reference (the reference
projection restarts pressure from zero, backend-cuda.cu:644).

Pinned by tests/test_unet_oracle.py: conv3 and the whole network against torch's
float64 conv3d / autograd, and the gradient against central finite differences.

Tensors are channels-last [z][y][x][c]; conv weights [27][cout][cin] with tap =
9*dz + 3*dy + dx and offsets (dx-1, dy-1, dz-1). `rnd` is applied to every stored
activation and activation gradient: identity for FP64, bf16_round to mirror the GPU's
BF16 tensors.
"""
import numpy as np

LAYERS = [(2, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 32), (32, 32), (32, 16), (16, 16), (16, 1)]


def param_count():
    return sum(27 * ci * co + co for ci, co in LAYERS)


def split_params(flat):
    """Public flat layout -> [(w [27][co][ci], b [co])] per layer."""
    out, p = [], 0
    for ci, co in LAYERS:
        w = np.asarray(flat[p:p + 27 * co * ci], dtype=np.float64).reshape(27, co, ci)
        p += 27 * co * ci
        b = np.asarray(flat[p:p + co], dtype=np.float64)
        p += co
        out.append((w, b))
    assert p == len(flat)
    return out


def join_params(layers):
    return np.concatenate([np.concatenate([w.reshape(-1), b.reshape(-1)]) for w, b in layers])


def bf16_round(a):
    """Round to the nearest BF16 value (ties to even), as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def identity(a):
    return np.asarray(a, dtype=np.float64)


def _shifted(xp, tap, shape):
    dz, dy, dx = tap // 9, (tap // 3) % 3, tap % 3
    z, y, x = shape
    return xp[dz:dz + z, dy:dy + y, dx:dx + x, :]


def conv3(x, w, b=None):
    """y[v][co] = sum_tap sum_ci w[tap][co][ci] x[v + off(tap)][ci] + b[co]."""
    z, y, xx, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((z, y, xx, w.shape[1]))
    for tap in range(27):
        out += _shifted(xp, tap, (z, y, xx)) @ w[tap].T
    if b is not None:
        out += b
    return out


def conv3_backward(x, w, g):
    """(dx, dw, db) of y = conv3(x, w, b) for upstream gradient g."""
    z, y, xx, ci = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dw = np.zeros_like(w)
    g2 = g.reshape(-1, g.shape[-1])
    for tap in range(27):
        dz, dy, dx = tap // 9, (tap // 3) % 3, tap % 3
        dw[tap] = g2.T @ _shifted(xp, tap, (z, y, xx)).reshape(-1, ci)
        dxp[dz:dz + z, dy:dy + y, dx:dx + xx, :] += g @ w[tap]
    return dxp[1:-1, 1:-1, 1:-1, :], dw, g2.sum(axis=0)


def pool2(x):
    z, y, xx, c = x.shape
    return x.reshape(z // 2, 2, y // 2, 2, xx // 2, 2, c).mean(axis=(1, 3, 5))


def up2(x):
    return x.repeat(2, axis=0).repeat(2, axis=1).repeat(2, axis=2)


def pool2_adjoint(g):
    return up2(g) / 8.0


def up2_adjoint(g):
    z, y, xx, c = g.shape
    return g.reshape(z // 2, 2, y // 2, 2, xx // 2, 2, c).sum(axis=(1, 3, 5))


def relu(a):
    return np.maximum(a, 0.0)


def forward(params, inp, rnd=identity):
    """inp [z][y][x][2] -> (q [z][y][x], cache)."""
    L = params
    a = {}
    a["a0"] = rnd(inp)
    a["a1"] = rnd(relu(conv3(a["a0"], *L[0])))
    a["a2"] = rnd(relu(conv3(a["a1"], *L[1])))
    a["p1"] = rnd(pool2(a["a2"]))
    a["a3"] = rnd(relu(conv3(a["p1"], *L[2])))
    a["a4"] = rnd(relu(conv3(a["a3"], *L[3])))
    a["p2"] = rnd(pool2(a["a4"]))
    a["a5"] = rnd(relu(conv3(a["p2"], *L[4])))
    a["a6"] = rnd(relu(conv3(a["a5"], *L[5])))
    a["u1"] = rnd(up2(a["a6"]))
    a["r7"] = rnd(relu(conv3(a["u1"], *L[6])))
    a["s7"] = rnd(a["r7"] + a["a4"])
    a["a8"] = rnd(relu(conv3(a["s7"], *L[7])))
    a["u2"] = rnd(up2(a["a8"]))
    a["r9"] = rnd(relu(conv3(a["u2"], *L[8])))
    a["s9"] = rnd(a["r9"] + a["a2"])
    a["a10"] = rnd(relu(conv3(a["s9"], *L[9])))
    q = conv3(a["a10"], *L[10])[..., 0]
    return q, a


def loss_and_grads(params, inp, target, rnd=identity):
    """MSE loss mean((q - t)^2) and its parameter gradient (list of (dw, db))."""
    q, a = forward(params, inp, rnd)
    n = q.size
    loss = float(np.mean((q - target) ** 2))
    gy = rnd((2.0 * (q - target) / n)[..., None])
    grads = [None] * len(LAYERS)
    mask = lambda g, m: rnd(g * (m > 0))

    dx, dw, db = conv3_backward(a["a10"], params[10][0], gy)
    grads[10] = (dw, db)
    g10 = mask(dx, a["a10"])
    dx, dw, db = conv3_backward(a["s9"], params[9][0], g10)
    grads[9] = (dw, db)
    ds9 = rnd(dx)
    g9 = mask(ds9, a["r9"])
    dx, dw, db = conv3_backward(a["u2"], params[8][0], g9)
    grads[8] = (dw, db)
    g8 = mask(rnd(up2_adjoint(rnd(dx))), a["a8"])
    dx, dw, db = conv3_backward(a["s7"], params[7][0], g8)
    grads[7] = (dw, db)
    ds7 = rnd(dx)
    g7 = mask(ds7, a["r7"])
    dx, dw, db = conv3_backward(a["u1"], params[6][0], g7)
    grads[6] = (dw, db)
    g6 = mask(rnd(up2_adjoint(rnd(dx))), a["a6"])
    dx, dw, db = conv3_backward(a["a5"], params[5][0], g6)
    grads[5] = (dw, db)
    g5 = mask(dx, a["a5"])
    dx, dw, db = conv3_backward(a["p2"], params[4][0], g5)
    grads[4] = (dw, db)
    t4 = rnd(pool2_adjoint(rnd(dx)))
    g4 = mask(ds7 + t4, a["a4"])
    dx, dw, db = conv3_backward(a["a3"], params[3][0], g4)
    grads[3] = (dw, db)
    g3 = mask(dx, a["a3"])
    dx, dw, db = conv3_backward(a["p1"], params[2][0], g3)
    grads[2] = (dw, db)
    t2 = rnd(pool2_adjoint(rnd(dx)))
    g2 = mask(ds9 + t2, a["a2"])
    dx, dw, db = conv3_backward(a["a1"], params[1][0], g2)
    grads[1] = (dw, db)
    g1 = mask(dx, a["a1"])
    _, dw, db = conv3_backward(a["a0"], params[0][0], g1)
    grads[0] = (dw, db)
    return loss, grads


def init_params(seed, scale=1.0):
    """He-normal weights (std sqrt(2 / (27 cin)) * scale), zero biases."""
    rng = np.random.default_rng(seed)
    layers = []
    for ci, co in LAYERS:
        w = rng.standard_normal((27, co, ci)) * np.sqrt(2.0 / (27 * ci)) * scale
        layers.append((w, np.zeros(co)))
    return layers


def adam_step(params, grads, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update in float64 (Kingma and Ba, "Adam: A Method for Stochastic
    Optimization", 2015, Algorithm 1), the update stable_fluids_unet_adam_step_cuda applies:
    m <- b1 m + (1 - b1) g, v <- b2 v + (1 - b2) g^2, then
    p <- p - lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps), t counting from 1.
    Returns (params, m, v) as new arrays."""
    g = np.asarray(grads, dtype=np.float64)
    m = beta1 * np.asarray(m, dtype=np.float64) + (1.0 - beta1) * g
    v = beta2 * np.asarray(v, dtype=np.float64) + (1.0 - beta2) * g * g
    m_hat = m / (1.0 - beta1 ** t)
    v_hat = v / (1.0 - beta2 ** t)
    p = np.asarray(params, dtype=np.float64) - lr * m_hat / (np.sqrt(v_hat) + eps)
    return p, m, v

