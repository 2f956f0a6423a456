"""TEST INFRASTRUCTURE — FP64 numpy oracle of the learned warm-start network.

Same layer order, layouts and arithmetic as iane_amd.learned, in float64:
channels-last volumes (z, y, x, c), weights (cout, 27, cin), tap = (dz*3+dy)*3+dx,
3x3x3 zero-padded cross-correlation, 2x average pooling, nearest upsampling
with additive skips, ReLU, prediction u + r, MSE loss. ``bf16=True`` rounds
the values the GPU stores in BF16 (inputs, activations, weights, the gradients
fed to the input-gradient convolution) to BF16 round-to-nearest-even, so a GPU
run differs from it only by FP32 accumulation order.

Pinned by tests/test_cnn_oracle.py: conv3d against the C oracle convolve with
flipped kernels (one layer = sum over input channels of the reference convolve
with the flipped weight, zero border), the analytic backward against central
finite differences, and the pooling / upsampling adjoint identities.
"""
from __future__ import annotations

import numpy as np

LAYERS = [("c1a", 1, 16, True), ("c1b", 16, 16, True), ("c2a", 16, 32, True), ("c3a", 32, 64, True),
          ("c3b", 64, 32, True), ("c2b", 32, 16, True), ("c1c", 16, 1, False)]


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to BF16 (round to nearest, ties to even) and return as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    bits = f.view(np.uint32).astype(np.uint64)
    rounded = ((bits + 0x7FFF + ((bits >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return rounded.view(np.float32).astype(np.float64)


def conv3d(x: np.ndarray, w: np.ndarray, b=None) -> np.ndarray:
    """x (nz, ny, nx, cin), w (cout, 27, cin) -> (nz, ny, nx, cout)."""
    nz, ny, nx, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((nz, ny, nx, w.shape[0]))
    for t in range(27):
        dz, dy, dx = t // 9, (t // 3) % 3, t % 3
        out += xp[dz:dz + nz, dy:dy + ny, dx:dx + nx, :] @ w[:, t, :].T
    if b is not None:
        out += b
    return out


def conv3d_input_grad(dy: np.ndarray, w: np.ndarray) -> np.ndarray:
    """Gradient of conv3d with respect to x: correlation of dy with w'[ci][t][co] = w[co][26-t][ci]."""
    wt = np.transpose(w[:, ::-1, :], (2, 1, 0))
    return conv3d(dy, wt)


def conv3d_weight_grad(x: np.ndarray, dy: np.ndarray):
    nz, ny, nx, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    dw = np.zeros((dy.shape[3], 27, x.shape[3]))
    for t in range(27):
        dz, dyy, dx = t // 9, (t // 3) % 3, t % 3
        window = xp[dz:dz + nz, dyy:dyy + ny, dx:dx + nx, :].reshape(-1, x.shape[3])
        dw[:, t, :] = dy.reshape(-1, dy.shape[3]).T @ window
    return dw, dy.reshape(-1, dy.shape[3]).sum(axis=0)


def pool2(x):
    nz, ny, nx, c = x.shape
    return x.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(1, 3, 5))


def pool2_backward(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 0), 2, 1), 2, 2) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up2_backward(dfine):
    nz, ny, nx, c = dfine.shape
    return dfine.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5))


def relu(x):
    return np.maximum(x, 0.0)


def forward(params: dict, u: np.ndarray, emulate_bf16: bool = False):
    """params[name] = (w, b); u (nz, ny, nx). Returns (r, activations)."""
    q = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    W = {n: (q(params[n][0]), np.asarray(params[n][1], dtype=np.float64)) for n, *_ in LAYERS}
    a = {"ubf": q(u[..., None])}
    a["a1a"] = q(relu(conv3d(a["ubf"], *W["c1a"])))
    a["a1b"] = q(relu(conv3d(a["a1a"], *W["c1b"])))
    a["p1"] = q(pool2(a["a1b"]))
    a["a2a"] = q(relu(conv3d(a["p1"], *W["c2a"])))
    a["p2"] = q(pool2(a["a2a"]))
    a["a3a"] = q(relu(conv3d(a["p2"], *W["c3a"])))
    a["a3b"] = q(relu(conv3d(a["a3a"], *W["c3b"])))
    a["s2"] = q(up2(a["a3b"]) + a["a2a"])
    a["a2b"] = q(relu(conv3d(a["s2"], *W["c2b"])))
    a["s1"] = q(up2(a["a2b"]) + a["a1b"])
    r = conv3d(a["s1"], *W["c1c"])[..., 0]
    return r, a, W


def loss(params, u, target):
    r, _, _ = forward(params, u)
    d = u + r - target
    return float(np.mean(d * d))


def backward(params: dict, u: np.ndarray, target: np.ndarray, emulate_bf16: bool = False):
    """Gradients of mean((u + r - target)^2) with respect to every (w, b)."""
    q = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    r, a, W = forward(params, u, emulate_bf16)
    n = u.size
    dr = (2.0 * (u + r - target) / n)[..., None]
    g = {}

    def wgrad(name, x, dy):
        g[name] = conv3d_weight_grad(x, dy)

    def dgrad(name, dy):
        return conv3d_input_grad(q(dy), W[name][0])

    def relu_back(y, dy):
        return np.where(y > 0, dy, 0.0)

    wgrad("c1c", a["s1"], dr)
    ds1 = dgrad("c1c", dr)
    d_a2b = relu_back(a["a2b"], up2_backward(ds1))
    d_a1b = ds1
    wgrad("c2b", a["s2"], d_a2b)
    ds2 = dgrad("c2b", d_a2b)
    d_a3b = relu_back(a["a3b"], up2_backward(ds2))
    d_a2a = ds2
    wgrad("c3b", a["a3a"], d_a3b)
    d_a3a = relu_back(a["a3a"], dgrad("c3b", d_a3b))
    wgrad("c3a", a["p2"], d_a3a)
    d_p2 = dgrad("c3a", d_a3a)
    d_a2a = relu_back(a["a2a"], d_a2a + pool2_backward(d_p2))
    wgrad("c2a", a["p1"], d_a2a)
    d_p1 = dgrad("c2a", d_a2a)
    d_a1b = relu_back(a["a1b"], d_a1b + pool2_backward(d_p1))
    wgrad("c1b", a["a1a"], d_a1b)
    d_a1a = relu_back(a["a1a"], dgrad("c1b", d_a1b))
    wgrad("c1a", a["ubf"], d_a1a)
    return r, g


def random_params(seed: int = 0, scale: float = 1.0):
    rng = np.random.default_rng(seed)
    params = {}
    for name, cin, cout, _ in LAYERS:
        w = rng.standard_normal((cout, 27, cin)) * np.sqrt(2.0 / (27 * cin)) * scale
        params[name] = (w, rng.standard_normal(cout) * 0.05)
    return params
