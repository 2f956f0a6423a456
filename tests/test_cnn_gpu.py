"""GPU parity of the warm-start network kernels against the FP64 oracle.

Single layers take BF16 inputs/weights on both sides, so they differ only by
FP32 accumulation: max|gpu - oracle| <= 1e-5 * max|oracle| (TOL_LAYER).
The whole network forward/backward is compared with the oracle emulating the
BF16 stores; a value that rounds to the other side of a BF16 step can move an
activation by 2^-8 relative, so the end-to-end tolerance is 2e-2 normwise (TOL_NET).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from iane_amd import _lib, learned
from oracle import cnn_oracle as cnn

pytestmark = pytest.mark.gpu
TOL_LAYER = 1e-5
TOL_NET = 2e-2


def dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return t.to(dtype) if dtype is not None else t


def bf(a):
    import torch
    return dev(np.asarray(a, np.float32), torch.bfloat16)


def rel(got, ref):
    got = np.asarray(got, np.float64)
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-30))


# (16, 32), (32, 64), (32, 16), (24, 40): whole load groups plus leftover K chunks;
# (64, 32): whole groups only; (8, 3), (8, 48): fewer chunks than one group;
# (12, 20): cin not a multiple of 8 (one chunk at a time); (1, 5): input layer
# with a partial channel group (channel-by-channel stores)
@pytest.mark.parametrize("cin,cout", [(1, 16), (16, 16), (16, 32), (32, 64), (64, 32), (32, 16), (16, 1), (8, 3),
                                      (24, 40), (8, 48), (12, 20), (1, 5)])
def test_conv3d_layer(cin, cout):
    import torch
    rng = np.random.default_rng(cin * 100 + cout)
    # odd y extent: the last row group of the generic kernel (4 rows) and the
    # last 8-row block of the others are partial
    shape = (6, 11, 37)
    x = cnn.bf16(rng.standard_normal(shape + (cin,)))
    w = cnn.bf16(rng.standard_normal((cout, 27, cin)) * 0.2)
    b = rng.standard_normal(cout).astype(np.float32)
    # device copies stay referenced until the stream is synchronised: a dropped
    # temporary would hand its memory to the next allocation before the launch runs
    xd, wd, bd = bf(x), bf(w), dev(b)
    out = torch.empty(shape + (cout,), dtype=torch.float32, device="cuda")
    _lib.call("iane_hip_cnn_conv3d", xd.data_ptr(), wd.data_ptr(), bd.data_ptr(), out.data_ptr(), 0, 0,
              shape[2], shape[1], shape[0], cin, cout, 0)
    ref = cnn.conv3d(x, w, b.astype(np.float64))
    assert rel(out.cpu().numpy(), ref) <= TOL_LAYER
    # ReLU + BF16 output
    outb = torch.empty(shape + (cout,), dtype=torch.bfloat16, device="cuda")
    _lib.call("iane_hip_cnn_conv3d", xd.data_ptr(), wd.data_ptr(), bd.data_ptr(), outb.data_ptr(), 1, 1,
              shape[2], shape[1], shape[0], cin, cout, 0)
    got = outb.float().cpu().numpy()
    np.testing.assert_allclose(got, cnn.bf16(np.maximum(ref, 0)), rtol=2 ** -7, atol=1e-6)


@pytest.mark.parametrize("cin,cout", [(16, 16), (32, 64), (16, 1), (1, 16)])
def test_input_and_weight_gradients(cin, cout):
    import torch
    rng = np.random.default_rng(cin + cout)
    shape = (4, 8, 20)
    x = cnn.bf16(rng.standard_normal(shape + (cin,)))
    w = cnn.bf16(rng.standard_normal((cout, 27, cin)) * 0.2)
    dy = rng.standard_normal(shape + (cout,)).astype(np.float32)
    # weight gradient (accumulating into zeroed buffers)
    xd, wd, dyd, dyb = bf(x), bf(w), dev(dy), bf(cnn.bf16(dy))
    dw = torch.zeros((cout, 27, cin), dtype=torch.float32, device="cuda")
    db = torch.zeros(cout, dtype=torch.float32, device="cuda")
    _lib.call("iane_hip_cnn_wgrad", xd.data_ptr(), dyd.data_ptr(), dw.data_ptr(), db.data_ptr(),
              shape[2], shape[1], shape[0], cin, cout, 0)
    ref_dw, ref_db = cnn.conv3d_weight_grad(x, dy.astype(np.float64))
    assert rel(dw.cpu().numpy(), ref_dw) <= TOL_LAYER
    assert rel(db.cpu().numpy(), ref_db) <= TOL_LAYER
    # input gradient = conv3d with the transposed flipped weights
    wt = torch.empty((cin, 27, cout), dtype=torch.bfloat16, device="cuda")
    _lib.call("iane_hip_cnn_transpose_weights", wd.data_ptr(), wt.data_ptr(), cin, cout, 0)
    dx = torch.empty(shape + (cin,), dtype=torch.float32, device="cuda")
    _lib.call("iane_hip_cnn_conv3d", dyb.data_ptr(), wt.data_ptr(), 0, dx.data_ptr(), 0, 0,
              shape[2], shape[1], shape[0], cout, cin, 0)
    assert rel(dx.cpu().numpy(), cnn.conv3d_input_grad(cnn.bf16(dy), w)) <= TOL_LAYER


def _net_with(params):
    import torch
    net = learned.warm_start_net(seed=1)
    for name, *_ in learned.LAYERS:
        net.w32[name].copy_(torch.from_numpy(params[name][0].astype(np.float32)))
        net.b32[name].copy_(torch.from_numpy(params[name][1].astype(np.float32)))
    net._refresh_bf16()
    return net


def test_network_forward_backward_vs_oracle():
    import torch
    rng = np.random.default_rng(7)
    params = cnn.random_params(3)
    params = {k: (cnn.bf16(w).astype(np.float32).astype(np.float64), b.astype(np.float32).astype(np.float64))
              for k, (w, b) in params.items()}
    shape = (8, 12, 16)
    u = (rng.random(shape) * 4).astype(np.float32)
    target = (u + rng.standard_normal(shape) * 0.5).astype(np.float32)
    net = _net_with(params)
    partials, used, grad = net.loss_and_grad(dev(u), dev(target))
    net.backward(grad)
    torch.cuda.synchronize()
    r_ref, g_ref = cnn.backward(params, u.astype(np.float64), target.astype(np.float64), emulate_bf16=True)
    assert rel(net._saved.r.cpu().numpy(), r_ref) <= TOL_NET
    for name, *_ in learned.LAYERS:
        assert rel(net.gw[name].cpu().numpy(), g_ref[name][0]) <= TOL_NET, name
        assert rel(net.gb[name].cpu().numpy(), g_ref[name][1]) <= TOL_NET, name
    loss = float(partials[:used].cpu().numpy().sum()) / u.size
    assert abs(loss - np.mean((u + r_ref - target) ** 2)) <= 1e-3 * loss


def test_training_reduces_the_loss():
    import torch
    rng = np.random.default_rng(9)
    net = learned.warm_start_net(seed=2)
    shape = (16, 16, 16)
    us = [dev((rng.random(shape) * 4).astype(np.float32)) for _ in range(2)]
    targets = [dev(np.asarray(u.cpu().numpy() * 1.5 + 0.2, np.float32)) for u in us]
    losses = []
    for step in range(60):
        i = step % 2
        partials, used, grad = net.loss_and_grad(us[i], targets[i])
        net.backward(grad)
        net.adam_step(3e-3)
        losses.append(float(partials[:used].cpu().numpy().sum()) / us[i].numel())
    torch.cuda.synchronize()
    assert np.mean(losses[-6:]) < 0.5 * np.mean(losses[:6])


def test_device_data_synthesis():
    import torch
    n = 200000
    lam = dev(np.full(n, 30.0, np.float32))
    out = torch.empty_like(lam)
    _lib.call("iane_hip_synth_poisson", lam.data_ptr(), out.data_ptr(), n, ctypes.c_uint(3), 0)
    v = out.cpu().numpy()
    assert abs(v.mean() - 30.0) < 0.1 and abs(v.var() - 30.0) < 1.0 and np.all(v == np.floor(v))
    lam2 = dev(np.full(n, 2.5, np.float32))
    _lib.call("iane_hip_synth_poisson", lam2.data_ptr(), out.data_ptr(), n, ctypes.c_uint(4), 0)
    v = out.cpu().numpy()
    assert abs(v.mean() - 2.5) < 0.03 and abs(v.var() - 2.5) < 0.1
    scratch = torch.empty(4097, dtype=torch.float32, device="cuda")
    x = dev(np.random.default_rng(0).random(123457).astype(np.float32))
    _lib.call("iane_hip_max", x.data_ptr(), x.numel(), scratch.data_ptr(), scratch.data_ptr() + 4096 * 4, 0)
    assert float(scratch[4096].item()) == float(x.max().item())
