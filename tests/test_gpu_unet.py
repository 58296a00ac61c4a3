"""The warm-start network's HIP kernels against the FP64 oracle (oracle/unet_oracle.py).

Bars:
  * one convolution (BF16 operands, FP32 accumulation and output) against FP64 on the
    same BF16 values: 1e-5 of max|y| (only the accumulation order differs);
  * weight and bias gradients of one layer: 1e-4 (FP32 atomics over the voxels);
  * the whole network (BF16 activations and activation gradients between layers, as the
    GPU stores them; the oracle rounds at the same points): loss, output and every
    layer's gradient within 1e-2 norm-wise (measured on MI355X: output 3.0e-3, worst
    layer gradient 1.8e-3); against unrounded FP64 the gap (up to 3.3e-2) is reported.
Grids include partial workgroup tiles (extents not multiples of 16 x 4 x 2).
"""
import numpy as np
import pytest

from conftest import sf
from oracle import unet_oracle as uo

pytestmark = pytest.mark.gpu
hip = sf.hip
U = sf.unet

PAIRS = [(16, 16), (16, 32), (32, 16), (32, 32), (32, 64), (64, 32), (64, 64)]


def bf16(a):
    return uo.bf16_round(a)


def dev(a):
    return hip.DeviceBuffer.from_array(np.ascontiguousarray(a))


@pytest.mark.parametrize("ci,co", PAIRS)
def test_conv3_forward_matches_fp64(ci, co):
    rng = np.random.default_rng(ci * 7 + co)
    shape = (5, 10, 37)
    x = bf16(rng.standard_normal(shape + (ci,)))
    w = bf16(rng.standard_normal((27, co, ci)) * 0.2)
    b = rng.standard_normal(co).astype(np.float32)
    y = hip.DeviceBuffer(np.prod(shape) * co * 4)
    xd, wd, bd = dev(U.to_bf16_bits(x)), dev(U.to_bf16_bits(w)), dev(b)
    for relu in (False, True):
        U.conv3(xd, wd, bd, y, ci, co, shape, relu=relu)
        hip.synchronize()
        got = hip.to_host(y.value, np.prod(shape) * co).reshape(shape + (co,))
        ref = uo.conv3(x, w, b.astype(np.float64))
        if relu:
            ref = np.maximum(ref, 0)
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print(f"conv3 {ci}->{co} relu={relu} err={err:.2e}", flush=True)
        assert err <= 1e-5


@pytest.mark.parametrize("ci,co", PAIRS)
def test_conv3_wgrad_matches_fp64(ci, co):
    rng = np.random.default_rng(ci * 11 + co)
    shape = (5, 10, 37)
    x = bf16(rng.standard_normal(shape + (ci,)))
    g = bf16(rng.standard_normal(shape + (co,)))
    dw = dev(np.zeros(27 * co * ci, np.float32))
    db = dev(np.zeros(co, np.float32))
    U.conv3_wgrad(dev(U.to_bf16_bits(x)), dev(U.to_bf16_bits(g)), dw, db, ci, co, shape)
    hip.synchronize()
    got_w = hip.to_host(dw.value, 27 * co * ci).reshape(27, co, ci)
    got_b = hip.to_host(db.value, co)
    _, ref_w, ref_b = uo.conv3_backward(x, np.zeros((27, co, ci)), g)
    ew = np.max(np.abs(got_w - ref_w)) / np.max(np.abs(ref_w))
    eb = np.max(np.abs(got_b - ref_b)) / np.max(np.abs(ref_b))
    print(f"wgrad {ci}->{co} dw err={ew:.2e} db err={eb:.2e}", flush=True)
    assert ew <= 1e-4 and eb <= 1e-4


def random_problem(shape, seed):
    rng = np.random.default_rng(seed)
    params = uo.init_params(seed)
    params = [(w, rng.standard_normal(b.shape) * 0.05) for w, b in params]
    inp = rng.standard_normal(shape + (2,))
    inp[..., 1] = (rng.random(shape) < 0.3).astype(np.float64)
    target = rng.standard_normal(shape)
    return params, inp, target


def test_network_gradient_matches_oracle():
    shape = (8, 12, 16)
    params, inp, target = random_problem(shape, 2)
    flat = uo.join_params(params).astype(np.float32)
    nz, ny, nx = shape
    with U.UNet(nx, ny, nz) as net:
        net.set_params(flat)
        assert np.array_equal(net.params(), flat)
        inp_d, tgt_d = dev(inp.astype(np.float32)), dev(target.astype(np.float32))
        out_d, loss_d = hip.DeviceBuffer(inp[..., 0].size * 4), dev(np.zeros(1, np.float32))
        net.forward(inp_d, out_d)
        net.backward(inp_d, tgt_d, loss_d)
        hip.synchronize()
        q = hip.to_host(out_d.value, inp[..., 0].size).reshape(shape)
        loss = float(hip.to_host(loss_d.value, 1)[0])
        grads = uo.split_params(net.grads().astype(np.float64))
    # the oracle with the GPU's BF16 weights and BF16 tensors between layers
    p_bf = [(bf16(w), np.asarray(b, np.float32).astype(np.float64)) for w, b in uo.split_params(flat.astype(np.float64))]
    q_ref, _ = uo.forward(p_bf, inp, uo.bf16_round)
    loss_ref, g_ref = uo.loss_and_grads(p_bf, inp, target, uo.bf16_round)
    _, g64 = uo.loss_and_grads(p_bf, inp, target)
    eq = np.max(np.abs(q - q_ref)) / np.max(np.abs(q_ref))
    el = abs(loss - loss_ref) / loss_ref
    print(f"network output err={eq:.2e} loss err={el:.2e}", flush=True)
    assert eq <= 1e-2 and el <= 1e-2
    for l, ((dw, db), (rw, rb), (fw, fb)) in enumerate(zip(grads, g_ref, g64)):
        ew = np.max(np.abs(dw - rw)) / np.max(np.abs(rw))
        eb = np.max(np.abs(db - rb)) / max(np.max(np.abs(rb)), 1e-30)
        e64 = np.max(np.abs(dw - fw)) / np.max(np.abs(fw))
        print(f"layer {l + 1}: dw err={ew:.2e} db err={eb:.2e} (vs unrounded fp64 {e64:.2e})", flush=True)
        assert ew <= 1e-2 and eb <= 1e-2, l


def test_adam_steps_match_the_oracle():
    """Three Adam steps on the GPU (FP32 moments and parameters) against the float64
    update of oracle/unet_oracle.adam_step, each fed the gradient the GPU computed: every
    parameter within 2 ulp of its value plus 1e-5 of the step size. eps = 1e-3 (about the
    gradients' size) pins its placement outside the square root; three steps pin both
    bias corrections."""
    shape = (8, 8, 8)
    rng = np.random.default_rng(11)
    inp = rng.standard_normal(shape + (2,)).astype(np.float32)
    target = rng.standard_normal(shape).astype(np.float32)
    nz, ny, nx = shape
    lr, b1, b2, eps = 2e-3, 0.9, 0.99, 1e-3
    with U.UNet(nx, ny, nz) as net:
        p = uo.join_params(uo.init_params(2)).astype(np.float32)
        net.set_params(p)
        ref, m, v = p.astype(np.float64), 0.0, 0.0
        inp_d, tgt_d = dev(inp), dev(target)
        for t in range(1, 4):
            net.backward(inp_d, tgt_d)
            hip.synchronize()
            g = net.grads().astype(np.float64)
            net.adam_step(lr, b1, b2, eps)
            hip.synchronize()
            got = net.params().astype(np.float64)
            ref, m, v = uo.adam_step(ref, g, m, v, t, lr, b1, b2, eps)
            err = np.abs(got - ref)
            bound = 2.0 ** -22 * np.abs(ref) + 1e-5 * lr
            print(f"adam step {t}: max err {err.max():.3e}, max step {np.max(np.abs(ref - p)):.3e}", flush=True)
            assert np.all(err <= bound), (t, float(np.max(err - bound)))
            # keep the reference on the GPU's rounded parameters (the next gradient is taken there)
            ref = got


def test_adam_training_reduces_loss():
    """A smooth target (box-filtered input channel 0) is learnable: 60 Adam steps on the
    network's own kernels cut the loss at least in half."""
    shape = (16, 16, 16)
    rng = np.random.default_rng(5)
    inp = np.zeros(shape + (2,))
    inp[..., 0] = rng.standard_normal(shape)
    target = sf.scenes.box_filter5(inp[..., 0]) * 4.0
    nz, ny, nx = shape
    with U.UNet(nx, ny, nz) as net:
        net.set_params(uo.join_params(uo.init_params(1)).astype(np.float32))
        inp_d, tgt_d = dev(inp.astype(np.float32)), dev(target.astype(np.float32))
        losses = []
        for it in range(60):
            loss_d = dev(np.zeros(1, np.float32))
            net.backward(inp_d, tgt_d, loss_d)
            net.adam_step(2e-3)
            hip.synchronize()
            losses.append(float(hip.to_host(loss_d.value, 1)[0]))
    print("losses", [round(v, 4) for v in losses[::10]], losses[-1], flush=True)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0]
