"""Learned warm start for Richardson–Lucy (no reference counterpart).

A 3-level 3-D encoder-decoder (16/32/64 channels, 3x3x3 convolutions, ReLU,
BF16 weights and activations, FP32 accumulation) predicts, from a blurred
Poisson volume y, the result of 200 RL iterations; RL then starts from that
prediction instead of max(y, eps) and stops once the Poisson log-likelihood
reaches that of plain 200-iteration RL.

    level 0 (full)  c1a 1->16, c1b 16->16 -------------------------+ skip
    level 1 (1/2)   pool, c2a 16->32 ------------------+ skip         |
    level 2 (1/4)   pool, c3a 32->64, c3b 64->32, up2 + skip, c2b 32->16, up2 + skip, c1c 16->1
    prediction = u + r,   u = y * (4 / max y)

Every array operation runs in libiane_hip.so (conv3d on MFMA, FP32 weight
gradients, pooling, upsampling, Adam); torch only holds device memory. The
training pairs are made on the GPU as well: blobs -> FFT blur -> Poisson noise
-> 200 iterations of the package's RL.
"""
from __future__ import annotations

import ctypes
import dataclasses
import pathlib

import numpy as np

from . import _lib, hip_policy

LAYERS = [
    # name, cin, cout, relu
    ("c1a", 1, 16, True), ("c1b", 16, 16, True), ("c2a", 16, 32, True), ("c3a", 32, 64, True),
    ("c3b", 64, 32, True), ("c2b", 32, 16, True), ("c1c", 16, 1, False),
]
WEIGHTS = pathlib.Path(__file__).resolve().parent / "weights" / "warm_start.npz"


def _call(name, *args):
    _lib.call(name, *args)


def _stream(policy):
    return (policy or hip_policy()).handle()


def _ptr(t) -> int:
    return t.data_ptr()


@dataclasses.dataclass
class _acts:
    u: object = None
    ubf: object = None
    a1a: object = None
    a1b: object = None
    p1: object = None
    a2a: object = None
    p2: object = None
    a3a: object = None
    a3b: object = None
    s2: object = None
    a2b: object = None
    s1: object = None
    r: object = None


class warm_start_net:
    """The network: FP32 master weights + BF16 copies, Adam state, forward/backward."""

    def __init__(self, seed: int = 5, policy=None):
        import torch
        from .synth import normal
        self.stream = _stream(policy)
        self.w32, self.b32, self.wbf, self.gw, self.gb = {}, {}, {}, {}, {}
        self.mw, self.vw, self.mb, self.vb = {}, {}, {}, {}
        self.step_count = 0
        for i, (name, cin, cout, _) in enumerate(LAYERS):
            fan_in = 27 * cin
            w = normal(seed, 100 + i, cout * 27 * cin).reshape(cout, 27, cin) * np.sqrt(2.0 / fan_in)
            if name == "c1c":
                w *= 0.1
            self.w32[name] = torch.from_numpy(w.astype(np.float32)).cuda()
            self.b32[name] = torch.zeros(cout, dtype=torch.float32, device="cuda")
            self.wbf[name] = torch.empty((cout, 27, cin), dtype=torch.bfloat16, device="cuda")
            for d in (self.gw, self.mw, self.vw):
                d[name] = torch.zeros_like(self.w32[name])
            for d in (self.gb, self.mb, self.vb):
                d[name] = torch.zeros_like(self.b32[name])
        self._refresh_bf16()

    # ------------------------------------------------------------ weights --
    def _refresh_bf16(self):
        for name, *_ in LAYERS:
            _call("iane_hip_cnn_to_bf16", _ptr(self.w32[name]), _ptr(self.wbf[name]), self.w32[name].numel(),
                  ctypes.c_float(1.0), self.stream)

    def parameter_count(self) -> int:
        return sum(self.w32[n].numel() + self.b32[n].numel() for n, *_ in LAYERS)

    def save(self, path=WEIGHTS):
        path = pathlib.Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        arrays = {}
        for name, *_ in LAYERS:
            arrays[f"{name}_w"] = self.w32[name].cpu().numpy()
            arrays[f"{name}_b"] = self.b32[name].cpu().numpy()
        arrays["step"] = np.int64(self.step_count)
        np.savez(path, **arrays)

    def load(self, path=WEIGHTS):
        import torch
        data = np.load(path, allow_pickle=False)
        for name, *_ in LAYERS:
            self.w32[name].copy_(torch.from_numpy(data[f"{name}_w"]))
            self.b32[name].copy_(torch.from_numpy(data[f"{name}_b"]))
        self.step_count = int(data["step"])
        self._refresh_bf16()
        return self

    # ------------------------------------------------------------ forward --
    def _conv(self, name, x, shape, out, out_bf16=True, relu=True):
        cin, cout = self.w32[name].shape[2], self.w32[name].shape[0]
        nz, ny, nx = shape
        _call("iane_hip_cnn_conv3d", _ptr(x), _ptr(self.wbf[name]), _ptr(self.b32[name]), _ptr(out),
              1 if out_bf16 else 0, 1 if relu else 0, nx, ny, nz, cin, cout, self.stream)
        return out

    def forward(self, u, keep: bool = False):
        """u: float32 (nz, ny, nx) normalised volume, extents divisible by 4.
        Returns r (float32 (nz, ny, nx)); prediction = u + r."""
        import torch
        nz, ny, nx = u.shape
        if nz % 4 or ny % 4 or nx % 4:
            raise _lib.invalid_argument("warm-start network needs extents divisible by 4")
        s0, s1, s2 = (nz, ny, nx), (nz // 2, ny // 2, nx // 2), (nz // 4, ny // 4, nx // 4)
        bf = torch.bfloat16
        a = _acts(u=u)
        a.ubf = torch.empty(s0 + (1,), dtype=bf, device="cuda")
        _call("iane_hip_cnn_to_bf16", _ptr(u), _ptr(a.ubf), u.numel(), ctypes.c_float(1.0), self.stream)
        a.a1a = self._conv("c1a", a.ubf, s0, torch.empty(s0 + (16,), dtype=bf, device="cuda"))
        a.a1b = self._conv("c1b", a.a1a, s0, torch.empty(s0 + (16,), dtype=bf, device="cuda"))
        a.p1 = torch.empty(s1 + (16,), dtype=bf, device="cuda")
        _call("iane_hip_cnn_pool2", _ptr(a.a1b), _ptr(a.p1), nx, ny, nz, 16, self.stream)
        a.a2a = self._conv("c2a", a.p1, s1, torch.empty(s1 + (32,), dtype=bf, device="cuda"))
        a.p2 = torch.empty(s2 + (32,), dtype=bf, device="cuda")
        _call("iane_hip_cnn_pool2", _ptr(a.a2a), _ptr(a.p2), nx // 2, ny // 2, nz // 2, 32, self.stream)
        a.a3a = self._conv("c3a", a.p2, s2, torch.empty(s2 + (64,), dtype=bf, device="cuda"))
        a.a3b = self._conv("c3b", a.a3a, s2, torch.empty(s2 + (32,), dtype=bf, device="cuda"))
        a.s2 = torch.empty(s1 + (32,), dtype=bf, device="cuda")
        _call("iane_hip_cnn_up2_add", _ptr(a.a3b), _ptr(a.a2a), _ptr(a.s2), nx // 2, ny // 2, nz // 2, 32, self.stream)
        a.a2b = self._conv("c2b", a.s2, s1, torch.empty(s1 + (16,), dtype=bf, device="cuda"))
        a.s1 = torch.empty(s0 + (16,), dtype=bf, device="cuda")
        _call("iane_hip_cnn_up2_add", _ptr(a.a2b), _ptr(a.a1b), _ptr(a.s1), nx, ny, nz, 16, self.stream)
        a.r = self._conv("c1c", a.s1, s0, torch.empty(s0, dtype=torch.float32, device="cuda"), out_bf16=False,
                         relu=False)
        if keep:
            self._saved = a
        return a.r

    # ----------------------------------------------------------- backward --
    def _wgrad(self, name, x, dy, shape):
        cin, cout = self.w32[name].shape[2], self.w32[name].shape[0]
        nz, ny, nx = shape
        _call("iane_hip_cnn_wgrad", _ptr(x), _ptr(dy), _ptr(self.gw[name]), _ptr(self.gb[name]), nx, ny, nz, cin, cout,
              self.stream)

    def _dgrad(self, name, dy, shape):
        """dx (f32, channels = cin) = conv3d(dy as bf16, transposed flipped weights)."""
        import torch
        cin, cout = self.w32[name].shape[2], self.w32[name].shape[0]
        nz, ny, nx = shape
        wt = torch.empty((cin, 27, cout), dtype=torch.bfloat16, device="cuda")
        _call("iane_hip_cnn_transpose_weights", _ptr(self.wbf[name]), _ptr(wt), cin, cout, self.stream)
        dyb = torch.empty(dy.shape, dtype=torch.bfloat16, device="cuda")
        _call("iane_hip_cnn_to_bf16", _ptr(dy), _ptr(dyb), dy.numel(), ctypes.c_float(1.0), self.stream)
        dx = torch.empty(tuple(shape) + (cin,), dtype=torch.float32, device="cuda")
        _call("iane_hip_cnn_conv3d", _ptr(dyb), _ptr(wt), 0, _ptr(dx), 0, 0, nx, ny, nz, cout, cin, self.stream)
        return dx

    def _relu_back(self, y, dy):
        _call("iane_hip_cnn_relu_backward", _ptr(y), _ptr(dy), dy.numel(), self.stream)
        return dy

    def backward(self, dr):
        """dr: float32 gradient of the loss with respect to r (= to the prediction)."""
        import torch
        a = self._saved
        nz, ny, nx = a.u.shape
        s0, s1, s2 = (nz, ny, nx), (nz // 2, ny // 2, nx // 2), (nz // 4, ny // 4, nx // 4)
        for name, *_ in LAYERS:
            self.gw[name].zero_()
            self.gb[name].zero_()
        dr4 = dr.reshape(s0 + (1,))
        self._wgrad("c1c", a.s1, dr4, s0)
        ds1 = self._dgrad("c1c", dr4, s0)                                   # (s0, 16)
        d_a2b = torch.empty(s1 + (16,), dtype=torch.float32, device="cuda")
        _call("iane_hip_cnn_up2_backward", _ptr(ds1), _ptr(d_a2b), nx, ny, nz, 16, self.stream)
        d_a1b = ds1                                                          # skip path (accumulated below)
        self._relu_back(a.a2b, d_a2b)
        self._wgrad("c2b", a.s2, d_a2b, s1)
        ds2 = self._dgrad("c2b", d_a2b, s1)                                  # (s1, 32)
        d_a3b = torch.empty(s2 + (32,), dtype=torch.float32, device="cuda")
        _call("iane_hip_cnn_up2_backward", _ptr(ds2), _ptr(d_a3b), nx // 2, ny // 2, nz // 2, 32, self.stream)
        d_a2a = ds2
        self._relu_back(a.a3b, d_a3b)
        self._wgrad("c3b", a.a3a, d_a3b, s2)
        d_a3a = self._dgrad("c3b", d_a3b, s2)                                # (s2, 64)
        self._relu_back(a.a3a, d_a3a)
        self._wgrad("c3a", a.p2, d_a3a, s2)
        d_p2 = self._dgrad("c3a", d_a3a, s2)                                 # (s2, 32)
        _call("iane_hip_cnn_pool2_backward", _ptr(d_p2), _ptr(d_a2a), nx // 2, ny // 2, nz // 2, 32, 1, self.stream)
        self._relu_back(a.a2a, d_a2a)
        self._wgrad("c2a", a.p1, d_a2a, s1)
        d_p1 = self._dgrad("c2a", d_a2a, s1)                                 # (s1, 16)
        _call("iane_hip_cnn_pool2_backward", _ptr(d_p1), _ptr(d_a1b), nx, ny, nz, 16, 1, self.stream)
        self._relu_back(a.a1b, d_a1b)
        self._wgrad("c1b", a.a1a, d_a1b, s0)
        d_a1a = self._dgrad("c1b", d_a1b, s0)                                # (s0, 16)
        self._relu_back(a.a1a, d_a1a)
        self._wgrad("c1a", a.ubf, d_a1a, s0)

    def adam_step(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        self.step_count += 1
        for name, *_ in LAYERS:
            _call("iane_hip_cnn_adam", _ptr(self.w32[name]), _ptr(self.gw[name]), _ptr(self.mw[name]),
                  _ptr(self.vw[name]), _ptr(self.wbf[name]), self.w32[name].numel(), ctypes.c_float(lr),
                  ctypes.c_float(beta1), ctypes.c_float(beta2), ctypes.c_float(eps), self.step_count, self.stream)
            _call("iane_hip_cnn_adam", _ptr(self.b32[name]), _ptr(self.gb[name]), _ptr(self.mb[name]),
                  _ptr(self.vb[name]), 0, self.b32[name].numel(), ctypes.c_float(lr), ctypes.c_float(beta1),
                  ctypes.c_float(beta2), ctypes.c_float(eps), self.step_count, self.stream)

    def loss_and_grad(self, u, target):
        """MSE of (u + r) against target; returns (loss, dr) and keeps activations."""
        import torch
        r = self.forward(u, keep=True)
        grad = torch.empty_like(r)
        partials = torch.zeros(1024, dtype=torch.float64, device="cuda")
        used = ctypes.c_int()
        _call("iane_hip_cnn_residual_mse", _ptr(u), _ptr(r), _ptr(target), 0, _ptr(grad), u.numel(), _ptr(partials),
              ctypes.byref(used), self.stream)
        return partials, used.value, grad


# ----------------------------------------------------------- normalisation --
def normalisation(y, policy=None):
    """scale = 4 / max(y) (device max reduction); returns (u = y * scale, scale)."""
    import torch
    stream = _stream(policy)
    scratch = torch.empty(4097, dtype=torch.float32, device="cuda")
    _call("iane_hip_max", _ptr(y), y.numel(), _ptr(scratch), _ptr(scratch) + 4096 * 4, stream)
    peak = float(scratch[4096].item())
    scale = 4.0 / max(peak, 1e-6)
    u = torch.empty_like(y)
    _lib.call("iane_hip_memcpy_async", _ptr(u), _ptr(y), y.numel() * 4, 3, stream)
    _call("iane_hip_scale", _ptr(u), u.numel(), ctypes.c_float(scale), stream)
    return u, scale


def warm_start(net: warm_start_net, y, epsilon: float = 1e-6, policy=None):
    """x0 = max((u + net(u)) / scale, eps) for an observed float32 volume y."""
    import torch
    stream = _stream(policy)
    u, scale = normalisation(y, policy)
    r = net.forward(u)
    x0 = torch.empty_like(y)
    _call("iane_hip_cnn_warm_start", _ptr(u), _ptr(r), _ptr(x0), y.numel(), ctypes.c_float(1.0 / scale),
          ctypes.c_float(epsilon), stream)
    return x0


# ------------------------------------------------------------ training data --
def make_pairs(count: int, shape=(128, 128, 128), iterations: int = 200, seed: int = 5, psf_sigma=(3.0, 1.5, 1.5),
               psf_size=(63, 31, 31), peak: float = 1000.0, blobs_per_voxel: float = 300 / (256 * 256 * 128),
               policy=None):
    """``count`` (observed, RL-`iterations` result) pairs made on the GPU: blobs
    (sigma U[1,4], amplitude U[0.1,1]) on 0.01 background -> blur by the PSF
    (FFT, zero border) -> scale to `peak` -> Poisson -> plain RL."""
    import torch
    from . import deconvolution, synth
    stream = _stream(policy)
    pol = hip_policy(stream)
    psf = torch.from_numpy(synth.gaussian_kernel(psf_sigma, psf_size).astype(np.float32)).cuda()
    plan = deconvolution.rl_plan(shape, psf, deconvolution.FFT, policy=pol)
    n = int(np.prod(shape))
    nblobs = max(4, int(round(blobs_per_voxel * n)))
    truth = torch.empty(shape, dtype=torch.float32, device="cuda")
    blurred = torch.empty_like(truth)
    scratch = torch.empty(4097, dtype=torch.float32, device="cuda")
    ys, xs = [], []
    for i in range(count):
        u = synth.uniform(seed * 100003 + i, 7, nblobs * 5).reshape(nblobs, 5)
        params = np.empty((nblobs, 5), np.float32)
        params[:, 0] = u[:, 0] * shape[0]
        params[:, 1] = u[:, 1] * shape[1]
        params[:, 2] = u[:, 2] * shape[2]
        params[:, 3] = 1.0 + 3.0 * u[:, 3]
        params[:, 4] = 0.1 + 0.9 * u[:, 4]
        dparams = torch.from_numpy(params).cuda()
        _call("iane_hip_synth_blobs", _ptr(truth), shape[2], shape[1], shape[0], ctypes.c_float(0.01), _ptr(dparams),
              nblobs, stream)
        _call("iane_hip_rl_plan_blur", plan._handle, _ptr(truth), _ptr(blurred), _ptr(plan._workspace), stream)
        _call("iane_hip_max", _ptr(blurred), n, _ptr(scratch), _ptr(scratch) + 4096 * 4, stream)
        top = float(scratch[4096].item())
        _call("iane_hip_scale", _ptr(blurred), n, ctypes.c_float(peak / max(top, 1e-12)), stream)
        y = torch.empty_like(truth)
        _call("iane_hip_synth_poisson", _ptr(blurred), _ptr(y), n, ctypes.c_uint(seed * 7919 + i), stream)
        x = torch.empty_like(truth)
        plan.run(y, x, iterations)
        ys.append(y)
        xs.append(x)
    _lib.call("iane_hip_stream_sync", stream)
    return ys, xs


def train(net: warm_start_net, ys, xs, steps: int, lr: float = 1e-3, log_every: int = 50, log=print, policy=None):
    """Adam on the MSE between u + r and the normalised RL result; one pair per step."""
    import torch
    stream = _stream(policy)
    history = []
    order = np.argsort(synth_uniform_order(len(ys), steps))
    for step in range(steps):
        i = int(order[step % len(order)]) % len(ys)
        u, scale = normalisation(ys[i], policy)
        target = torch.empty_like(u)
        _lib.call("iane_hip_memcpy_async", _ptr(target), _ptr(xs[i]), xs[i].numel() * 4, 3, stream)
        _call("iane_hip_scale", _ptr(target), target.numel(), ctypes.c_float(scale), stream)
        partials, used, grad = net.loss_and_grad(u, target)
        net.backward(grad)
        net.adam_step(lr)
        if step % log_every == 0 or step == steps - 1:
            loss = float(partials[:used].cpu().numpy().sum()) / u.numel()
            history.append((step, loss))
            log(f"[train] step {step} mse {loss:.5f}")
    return history


def synth_uniform_order(count: int, steps: int):
    from .synth import uniform
    return uniform(97, 3, max(count, steps))
