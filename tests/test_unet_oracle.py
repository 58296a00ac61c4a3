"""The FP64 network oracle (oracle/unet_oracle.py) pinned on the CPU.

* conv3 equals torch's float64 conv3d (cross-correlation, zero padding 1) with the
  weight layout [27][cout][cin] <-> torch [cout][cin][dz][dy][dx];
* the whole network and its hand-derived gradient equal a torch float64 build of the
  same graph (conv3d, avg_pool3d, nearest interpolate, ReLU, additive skips) under
  autograd;
* the gradient matches central finite differences of the loss.
A dropped skip, a wrong tap order, a transposed weight or a missing ReLU mask fails one
of these.
"""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import unet_oracle as uo


def to_torch_w(w):
    co, ci = w.shape[1], w.shape[2]
    return torch.from_numpy(w.reshape(3, 3, 3, co, ci).transpose(3, 4, 0, 1, 2).copy())


def to_torch_x(x):
    return torch.from_numpy(np.ascontiguousarray(x.transpose(3, 0, 1, 2)))[None]


def from_torch_x(t):
    return t[0].detach().numpy().transpose(1, 2, 3, 0)


@pytest.mark.parametrize("shape,ci,co", [((4, 6, 5), 3, 2), ((8, 8, 8), 16, 32), ((3, 5, 7), 64, 16)])
def test_conv3_matches_torch_conv3d(shape, ci, co):
    rng = np.random.default_rng(ci * 100 + co)
    x = rng.standard_normal(shape + (ci,))
    w = rng.standard_normal((27, co, ci))
    b = rng.standard_normal(co)
    ours = uo.conv3(x, w, b)
    ref = from_torch_x(F.conv3d(to_torch_x(x), to_torch_w(w), torch.from_numpy(b), padding=1))
    assert np.max(np.abs(ours - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_conv3_tap_convention():
    """A single 1 at tap (dz,dy,dx) = (2,1,0) reads x[z+1][y][x-1]."""
    x = np.zeros((5, 5, 5, 1))
    x[3, 2, 1, 0] = 1.0
    w = np.zeros((27, 1, 1))
    w[9 * 2 + 3 * 1 + 0, 0, 0] = 1.0
    y = uo.conv3(x, w)
    assert y[2, 2, 2, 0] == 1.0 and np.sum(y) == 1.0


def torch_net(params, inp):
    P = [(to_torch_w(w).requires_grad_(), torch.from_numpy(b).requires_grad_()) for w, b in params]
    c = lambda t, l: F.conv3d(t, P[l][0], P[l][1], padding=1)
    a0 = to_torch_x(inp)
    a1 = F.relu(c(a0, 0))
    a2 = F.relu(c(a1, 1))
    a3 = F.relu(c(F.avg_pool3d(a2, 2), 2))
    a4 = F.relu(c(a3, 3))
    a5 = F.relu(c(F.avg_pool3d(a4, 2), 4))
    a6 = F.relu(c(a5, 5))
    s7 = F.relu(c(F.interpolate(a6, scale_factor=2, mode="nearest"), 6)) + a4
    a8 = F.relu(c(s7, 7))
    s9 = F.relu(c(F.interpolate(a8, scale_factor=2, mode="nearest"), 8)) + a2
    a10 = F.relu(c(s9, 9))
    q = c(a10, 10)[0, 0]
    return q, P


def test_network_and_gradient_match_torch_autograd():
    rng = np.random.default_rng(3)
    params = uo.init_params(7)
    params = [(w, rng.standard_normal(b.shape) * 0.1) for w, b in params]
    inp = rng.standard_normal((8, 12, 8, 2))
    target = rng.standard_normal((8, 12, 8))
    loss, grads = uo.loss_and_grads(params, inp, target)
    q_t, P = torch_net(params, inp)
    loss_t = torch.mean((q_t - torch.from_numpy(target)) ** 2)
    loss_t.backward()
    assert abs(loss - loss_t.item()) <= 1e-12 * abs(loss_t.item())
    for l, ((dw, db), (tw, tb)) in enumerate(zip(grads, P)):
        ref_w = tw.grad.numpy().transpose(2, 3, 4, 0, 1).reshape(27, dw.shape[1], dw.shape[2])
        scale = max(np.max(np.abs(ref_w)), 1e-30)
        assert np.max(np.abs(dw - ref_w)) <= 1e-10 * scale, f"layer {l} weights"
        assert np.max(np.abs(db - tb.grad.numpy())) <= 1e-10 * max(np.max(np.abs(tb.grad.numpy())), 1e-30), f"layer {l} bias"


def test_gradient_central_differences():
    rng = np.random.default_rng(11)
    params = uo.init_params(5)
    params = [(w, rng.standard_normal(b.shape) * 0.1) for w, b in params]
    inp = rng.standard_normal((8, 8, 8, 2))
    target = rng.standard_normal((8, 8, 8))
    _, grads = uo.loss_and_grads(params, inp, target)
    flat = uo.join_params(params)
    gflat = uo.join_params(grads)
    h = 1e-6
    for idx in rng.choice(flat.size, size=24, replace=False):
        p1, p2 = flat.copy(), flat.copy()
        p1[idx] += h
        p2[idx] -= h
        fd = (uo.loss_and_grads(uo.split_params(p1), inp, target)[0] - uo.loss_and_grads(uo.split_params(p2), inp, target)[0]) / (2 * h)
        assert abs(fd - gflat[idx]) <= 1e-6 + 1e-4 * abs(fd), (idx, fd, gflat[idx])


def test_param_count_and_layout():
    assert uo.param_count() == 319248 + 321
    flat = np.arange(uo.param_count(), dtype=np.float64)
    assert np.array_equal(uo.join_params(uo.split_params(flat)), flat)


def test_bf16_round():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0, 3.0e38], dtype=np.float64)
    r = uo.bf16_round(vals)
    assert r[0] == 1.0 and r[1] == 1.0        # tie to even
    assert r[2] == 1.0 + 2 ** -7               # above the tie rounds up
    assert r[3] == -2.5 and r[4] == 0.0
    assert np.isfinite(r[5])


def test_adam_step_matches_torch_adam():
    """The oracle's Adam update equals torch.optim.Adam (float64, no weight decay, no
    amsgrad) over 6 steps with changing gradients, eps large enough to pin its placement."""
    rng = np.random.default_rng(3)
    p0 = rng.standard_normal(257)
    grads = [rng.standard_normal(257) * 10.0 ** rng.uniform(-4, 0) for _ in range(6)]
    lr, b1, b2, eps = 3e-3, 0.85, 0.99, 1e-3
    tp = torch.tensor(p0.copy(), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=lr, betas=(b1, b2), eps=eps)
    p, m, v = p0.copy(), np.zeros_like(p0), np.zeros_like(p0)
    for t, g in enumerate(grads, start=1):
        opt.zero_grad()
        tp.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        p, m, v = uo.adam_step(p, g, m, v, t, lr, b1, b2, eps)
        np.testing.assert_allclose(p, tp.detach().numpy(), rtol=0, atol=1e-15)


def test_adam_step_closed_forms():
    """t = 1 from zero moments: the step is lr g / (|g| + eps) (bias corrections cancel the
    (1 - b) factors); a zero gradient leaves the parameter and decays the moments."""
    g = np.array([2.0, -0.5, 0.0, 1e-3])
    p, m, v = uo.adam_step(np.ones(4), g, np.zeros(4), np.zeros(4), 1, 0.1, 0.9, 0.999, 1e-3)
    np.testing.assert_allclose(p, 1.0 - 0.1 * g / (np.abs(g) + 1e-3), rtol=1e-14, atol=0)
    np.testing.assert_allclose(m, 0.1 * g, rtol=1e-15)
    np.testing.assert_allclose(v, 1e-3 * g * g, rtol=1e-12)
    p2, m2, v2 = uo.adam_step(p, np.zeros(4), m, v, 2, 0.1, 0.9, 0.999, 1e-3)
    np.testing.assert_allclose(m2, 0.9 * m, rtol=1e-15)
    np.testing.assert_allclose(v2, 0.999 * v, rtol=1e-15)
    # second step: m_hat = 0.09 g / 0.19, v_hat = 0.000999 g^2 / 0.001999
    expect = p - 0.1 * (0.09 * g / 0.19) / (np.sqrt(0.000999 * g * g / 0.001999) + 1e-3)
    np.testing.assert_allclose(p2, expect, rtol=1e-14, atol=0)
