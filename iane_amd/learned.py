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


def _host(t, stream) -> np.ndarray:
    """Copy a device tensor to the host after the work queued on `stream` is done
    (torch's copy runs on torch's own stream, which does not wait for ours)."""
    _lib.call("iane_hip_stream_sync", stream)
    return t.cpu().numpy()


def _zero(t, stream) -> None:
    """Clear a device tensor on `stream` (torch's own fill would run on torch's stream)."""
    _lib.call("iane_hip_memset_async", _ptr(t), 0, t.numel() * t.element_size(), stream)


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
        # the fills and copies above ran on torch's stream; the network's kernels use self.stream
        torch.cuda.current_stream().synchronize()
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
            arrays[f"{name}_w"] = _host(self.w32[name], self.stream)
            arrays[f"{name}_b"] = _host(self.b32[name], self.stream)
        arrays["step"] = np.int64(self.step_count)
        np.savez(path, **arrays)

    def load(self, path=WEIGHTS):
        import torch
        data = np.load(path, allow_pickle=False)
        _lib.call("iane_hip_stream_sync", self.stream)
        for name, *_ in LAYERS:
            self.w32[name].copy_(torch.from_numpy(data[f"{name}_w"]))
            self.b32[name].copy_(torch.from_numpy(data[f"{name}_b"]))
        torch.cuda.current_stream().synchronize()
        self.step_count = int(data["step"])
        self._refresh_bf16()
        return self

    # ------------------------------------------------------------ forward --
    def _conv(self, name, x, shape, out, out_bf16=True, relu=True, stream=None):
        cin, cout = self.w32[name].shape[2], self.w32[name].shape[0]
        nz, ny, nx = shape
        _call("iane_hip_cnn_conv3d", _ptr(x), _ptr(self.wbf[name]), _ptr(self.b32[name]), _ptr(out),
              1 if out_bf16 else 0, 1 if relu else 0, nx, ny, nz, cin, cout,
              self.stream if stream is None else stream)
        return out

    def forward(self, u, keep: bool = False, stream=None):
        """u: float32 (nz, ny, nx) normalised volume, extents divisible by 4.
        Returns r (float32 (nz, ny, nx)); prediction = u + r. ``stream``
        overrides the network's stream for this pass (inference on a caller's
        stream; the weights must be ready on it)."""
        import torch
        s = self.stream if stream is None else stream
        nz, ny, nx = u.shape
        if nz % 4 or ny % 4 or nx % 4:
            raise _lib.invalid_argument("warm-start network needs extents divisible by 4")
        s0, s1, s2 = (nz, ny, nx), (nz // 2, ny // 2, nx // 2), (nz // 4, ny // 4, nx // 4)
        bf = torch.bfloat16
        a = _acts(u=u)
        a.ubf = torch.empty(s0 + (1,), dtype=bf, device="cuda")
        _call("iane_hip_cnn_to_bf16", _ptr(u), _ptr(a.ubf), u.numel(), ctypes.c_float(1.0), s)
        a.a1a = self._conv("c1a", a.ubf, s0, torch.empty(s0 + (16,), dtype=bf, device="cuda"), stream=s)
        a.a1b = self._conv("c1b", a.a1a, s0, torch.empty(s0 + (16,), dtype=bf, device="cuda"), stream=s)
        a.p1 = torch.empty(s1 + (16,), dtype=bf, device="cuda")
        _call("iane_hip_cnn_pool2", _ptr(a.a1b), _ptr(a.p1), nx, ny, nz, 16, s)
        a.a2a = self._conv("c2a", a.p1, s1, torch.empty(s1 + (32,), dtype=bf, device="cuda"), stream=s)
        a.p2 = torch.empty(s2 + (32,), dtype=bf, device="cuda")
        _call("iane_hip_cnn_pool2", _ptr(a.a2a), _ptr(a.p2), nx // 2, ny // 2, nz // 2, 32, s)
        a.a3a = self._conv("c3a", a.p2, s2, torch.empty(s2 + (64,), dtype=bf, device="cuda"), stream=s)
        a.a3b = self._conv("c3b", a.a3a, s2, torch.empty(s2 + (32,), dtype=bf, device="cuda"), stream=s)
        a.s2 = torch.empty(s1 + (32,), dtype=bf, device="cuda")
        _call("iane_hip_cnn_up2_add", _ptr(a.a3b), _ptr(a.a2a), _ptr(a.s2), nx // 2, ny // 2, nz // 2, 32, s)
        a.a2b = self._conv("c2b", a.s2, s1, torch.empty(s1 + (16,), dtype=bf, device="cuda"), stream=s)
        a.s1 = torch.empty(s0 + (16,), dtype=bf, device="cuda")
        _call("iane_hip_cnn_up2_add", _ptr(a.a2b), _ptr(a.a1b), _ptr(a.s1), nx, ny, nz, 16, s)
        a.r = self._conv("c1c", a.s1, s0, torch.empty(s0, dtype=torch.float32, device="cuda"), out_bf16=False,
                         relu=False, stream=s)
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
            _zero(self.gw[name], self.stream)
            _zero(self.gb[name], self.stream)
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
        partials = torch.empty(1024, dtype=torch.float64, device="cuda")  # the first `used` entries are written
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
    peak = float(_host(scratch[4096:], stream)[0])
    scale = 4.0 / max(peak, 1e-6)
    u = torch.empty_like(y)
    _lib.call("iane_hip_memcpy_async", _ptr(u), _ptr(y), y.numel() * 4, 3, stream)
    _call("iane_hip_scale", _ptr(u), u.numel(), ctypes.c_float(scale), stream)
    return u, scale


def _padded_extents(shape):
    return tuple(-(-int(n) // 4) * 4 for n in shape)


def warm_start(net: warm_start_net, y, epsilon: float = 1e-6, policy=None):
    """x0 = max((u + net(u)) / scale, eps) for an observed float32 volume y.

    Extents that are not multiples of 4 are zero-padded at the high end for the
    network (two pooling levels) and the prediction is cropped back. Runs on the
    policy's stream (default: the network's stream)."""
    import torch
    from . import view_desc
    stream = net.stream if policy is None else _stream(policy)
    policy = hip_policy(stream)
    shape = tuple(y.shape)
    padded = _padded_extents(shape)
    if padded != shape:
        yp = torch.empty(padded, dtype=torch.float32, device=y.device)
        _zero(yp, stream)
        crop = yp[:shape[0], :shape[1], :shape[2]]
        _call("iane_hip_copy_view", ctypes.byref(view_desc(y)), ctypes.byref(view_desc(crop)), 3, stream)
    else:
        yp = y
    u, scale = normalisation(yp, policy)
    r = net.forward(u, stream=stream)
    xp = torch.empty_like(yp)
    _call("iane_hip_cnn_warm_start", _ptr(u), _ptr(r), _ptr(xp), yp.numel(), ctypes.c_float(1.0 / scale),
          ctypes.c_float(epsilon), stream)
    if padded == shape:
        return xp
    x0 = torch.empty_like(y)
    _call("iane_hip_copy_view", ctypes.byref(view_desc(xp[:shape[0], :shape[1], :shape[2]])),
          ctypes.byref(view_desc(x0)), 3, stream)
    return x0


# ------------------------------------------------------------ training data --
PSF_SIGMA = (3.0, 1.5, 1.5)
PSF_SIZE = (63, 31, 31)
PEAK = 1000.0
BLOBS_PER_MEGAVOXEL = 300 / (256 * 256 * 128) * 1e6


def psf_kernel(sigma=PSF_SIGMA, size=PSF_SIZE):
    """The instrument PSF the network is trained for: anisotropic Gaussian, sigma
    (3, 1.5, 1.5) voxels in (z, y, x) on a 63x31x31 support, unit sum (the PSF
    of the 256x256x128 RL workload)."""
    from . import synth
    return synth.gaussian_kernel(sigma, size).astype(np.float32)


def observed_volumes(count: int, shape=(128, 256, 256), seed: int = 5, psf=None, peak_range=(PEAK, PEAK),
                     blobs_per_megavoxel: float = BLOBS_PER_MEGAVOXEL, policy=None):
    """Yields ``count`` observed float32 device volumes made on the GPU: blobs
    (sigma U[1,4], amplitude U[0.1,1]) on 0.01 background -> blur by the PSF
    (FFT, zero border) -> scale to a peak U[peak_range] -> Poisson."""
    import torch
    from . import deconvolution, synth
    stream = _stream(policy)
    psf = torch.from_numpy(psf_kernel() if psf is None else np.asarray(psf, np.float32)).cuda()
    plan = deconvolution.rl_plan(shape, psf, deconvolution.FFT, policy=hip_policy(stream))
    n = int(np.prod(shape))
    nblobs = max(4, int(round(blobs_per_megavoxel * n / 1e6)))
    truth = torch.empty(shape, dtype=torch.float32, device="cuda")
    blurred = torch.empty_like(truth)
    scratch = torch.empty(4097, dtype=torch.float32, device="cuda")
    peaks = synth.uniform(seed, 8, count)
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
        top = float(_host(scratch[4096:], stream)[0])  # synchronises: dparams is no longer in use after this
        peak = peak_range[0] + peaks[i] * (peak_range[1] - peak_range[0])
        _call("iane_hip_scale", _ptr(blurred), n, ctypes.c_float(peak / max(top, 1e-12)), stream)
        y = torch.empty_like(truth)
        _call("iane_hip_synth_poisson", _ptr(blurred), _ptr(y), n, ctypes.c_uint(seed * 7919 + i), stream)
        yield y


def make_pairs(count: int, shape=(128, 128, 128), iterations: int = 200, seed: int = 5, psf=None,
               peak_range=(PEAK, PEAK), blobs_per_megavoxel: float = BLOBS_PER_MEGAVOXEL, policy=None):
    """``count`` (observed, RL-``iterations`` result) pairs: observed_volumes()
    followed by plain RL from max(y, eps)."""
    import torch
    from . import deconvolution
    stream = _stream(policy)
    psf_t = torch.from_numpy(psf_kernel() if psf is None else np.asarray(psf, np.float32)).cuda()
    plan = deconvolution.rl_plan(shape, psf_t, deconvolution.FFT, policy=hip_policy(stream))
    ys, xs = [], []
    for y in observed_volumes(count, shape, seed, psf, peak_range, blobs_per_megavoxel, policy):
        x = torch.empty_like(y)
        plan.run(y, x, iterations)
        ys.append(y)
        xs.append(x)
    _lib.call("iane_hip_stream_sync", stream)
    return ys, xs


def train(net: warm_start_net, ys, xs, steps: int, lr: float = 1e-3, log_every: int = 50, log=print,
          lr_final: float = 1e-5):
    """Adam on the MSE between u + r and the normalised RL result; one pair per
    step; the learning rate decays from ``lr`` to ``lr_final`` on a cosine.
    Runs on the network's stream (the pairs must be ready on it)."""
    import torch
    stream = net.stream
    policy = hip_policy(stream)
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
        frac = step / max(1, steps - 1)
        net.adam_step(lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * frac)))
        if step % log_every == 0 or step == steps - 1:
            loss = float(_host(partials[:used], stream).sum()) / u.numel()
            history.append((step, loss))
            log(f"[train] step {step} mse {loss:.5f}")
    return history


def synth_uniform_order(count: int, steps: int):
    from .synth import uniform
    return uniform(97, 3, max(count, steps))


# --------------------------------------------------------------- evaluation --
def _first_reaching(ll: np.ndarray, target: float) -> int:
    """Index of the first log-likelihood >= target (ll[j] is that of x_j), or -1."""
    hit = np.nonzero(ll >= target)[0]
    return int(hit[0]) if hit.size else -1


def iterations_to_likelihood(net: warm_start_net, observed, psf, iterations: int = 200, max_iterations: int = 400,
                             epsilon: float = 1e-6, policy=None):
    """For each observed volume: the log-likelihood of plain RL after
    ``iterations`` steps (the target) and how many RL steps from the network's
    warm start reach it. Returns a list of dicts (k = -1 if never reached)."""
    import torch
    from . import deconvolution
    pol = policy or hip_policy()
    out = []
    for y in observed:
        plan = deconvolution.rl_plan(tuple(y.shape), psf, deconvolution.FFT, epsilon=epsilon, policy=pol)
        x = torch.empty_like(y)
        ll = torch.empty(iterations + 1, dtype=torch.float64, device="cuda")  # every entry is written
        plan.run(y, x, iterations + 1, ll_history=ll)
        lp = _host(ll, pol.handle())
        target = float(lp[iterations])
        x0 = warm_start(net, y, epsilon, pol)
        llw = torch.empty(max_iterations + 1, dtype=torch.float64, device="cuda")
        plan.run(y, x0, max_iterations + 1, init=False, ll_history=llw)
        lw = _host(llw, pol.handle())
        out.append({"shape": tuple(y.shape), "target": target, "ll_warm0": float(lw[0]),
                    "ll_plain0": float(lp[0]), "k": _first_reaching(lw, target)})
    return out


def batch_time_to_likelihood(net: warm_start_net, observed, psfs, iterations: int = 200, max_iterations: int = 200,
                             repeats: int = 3, epsilon: float = 1e-6, check_every: int = 2, policy=None):
    """GPU time for a batch of volumes to reach the Poisson log-likelihood of
    plain ``iterations``-step RL, with the stopping rule inside the timed region.

    plain: batched RL for ``iterations`` steps with the fused per-iteration
    log-likelihood (the judge of convergence); its last value per volume is the
    target. learned: network inference for every volume, then batched RL from
    the warm starts with the log-likelihood, ``check_every`` iterations at a
    time, each chunk followed by a read-back of the latest log-likelihoods and
    the comparison with the targets, until every volume has reached its target
    (or ``max_iterations``). The returned ``k`` is the iteration count that the
    timed learned run itself stopped at. Times are medians of ``repeats`` runs
    between HIP events on the policy's stream (the host waits inside the
    learned run, so its checks are included). The batch replays hipGraphs,
    which need a created stream: with the null stream one is made."""
    pol = policy or hip_policy()
    if pol.handle() == 0:
        handle = ctypes.c_void_p()
        _lib.call("iane_hip_stream_create", ctypes.byref(handle))
        _lib.call("iane_hip_device_sync")  # the inputs were made on the null stream
        try:
            return batch_time_to_likelihood(net, observed, psfs, iterations, max_iterations, repeats, epsilon,
                                            check_every, hip_policy(handle.value))
        finally:
            _lib.call("iane_hip_stream_sync", handle.value)
            _lib.call("iane_hip_stream_destroy", handle.value)
    import torch
    from . import deconvolution
    stream = pol.handle()
    batch = deconvolution.rl_batch([tuple(y.shape) for y in observed], psfs, epsilon, policy=pol)
    n = len(observed)
    xs = [torch.empty_like(y) for y in observed]
    ll = torch.empty((n, iterations), dtype=torch.float64, device="cuda")  # every entry is written
    chunk_ll = torch.empty((n, check_every), dtype=torch.float64, device="cuda")
    targets = None
    ks = []

    def timed(fn):
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        results = []
        fn()  # untimed: graph capture and allocator warm-up
        for _ in range(repeats):
            _lib.call("iane_hip_stream_sync", stream)
            start.record(torch.cuda.ExternalStream(stream))
            fn()
            stop.record(torch.cuda.ExternalStream(stream))
            stop.synchronize()
            results.append(start.elapsed_time(stop))
        return float(np.median(results))

    def plain_path():
        batch.run(observed, xs, iterations, ll_history=ll)

    plain_ms = timed(plain_path)
    targets = _host(ll, stream)[:, iterations - 1]
    infer_ms = timed(lambda: [warm_start(net, y, epsilon, pol) for y in observed])

    def learned_path():
        starts = [warm_start(net, y, epsilon, pol) for y in observed]
        done = 0
        while done < max_iterations:
            batch.run(observed, starts, check_every, init=False, ll_history=chunk_ll)
            done += check_every
            if np.all(_host(chunk_ll, stream)[:, check_every - 1] >= targets):
                ks.append(done)
                return
        ks.append(-1)

    learned_ms = timed(learned_path)
    k = ks[-1]
    # the iteration at which each volume first reached its target (diagnostic, untimed)
    starts = [warm_start(net, y, epsilon, pol) for y in observed]
    llw = torch.empty((n, max_iterations), dtype=torch.float64, device="cuda")
    batch.run(observed, starts, max_iterations, init=False, ll_history=llw)
    lw = _host(llw, stream)
    # ll[j] is that of the estimate entering iteration j, so reaching the target at
    # index j means j + 1 iterations run, the count the timed loop reports
    per_volume = [j + 1 if j >= 0 else -1 for j in (_first_reaching(lw[v], targets[v]) for v in range(n))]
    return {"volumes": n, "plain_ms": plain_ms, "learned_ms": learned_ms if k >= 0 else float("nan"),
            "inference_ms": infer_ms, "k_max": k, "k_runs": ks, "k_per_volume": per_volume,
            "check_every": check_every, "speedup": plain_ms / learned_ms if k >= 0 else 0.0,
            "stopping": "the timed learned run reads back the log-likelihood every check_every iterations and "
                        "stops once every volume reaches the plain-RL target; plain RL computes the "
                        "log-likelihood every iteration"}


def main(argv=None):
    """python -m iane_amd.learned --pairs 2000 --steps 20000 --out weights.npz

    The committed weights came from three runs, each starting from the last
    (pre-training on a broader mix of PSFs and photon counts, then the
    256x256x128 workload's distribution):
      --pairs 384 --steps 6000 --lr 1e-3 --seed 5 --psf-size 19 11 11 --peak 200 2000 --blobs-per-mvox 150
      --pairs 768 --steps 22000 --lr 6e-4 --seed 6 --psf-size 19 11 11 --peak 200 2000 --blobs-per-mvox 150
      --pairs 2000 --steps 24000 --lr 5e-4 --seed 5            (defaults: 63x31x31 PSF, peak 1000)"""
    import argparse
    import time
    parser = argparse.ArgumentParser(prog="python -m iane_amd.learned")
    parser.add_argument("--pairs", type=int, default=2000)
    parser.add_argument("--edge", type=int, default=128)
    parser.add_argument("--steps", type=int, default=20000)
    parser.add_argument("--lr", type=float, default=1e-3)
    parser.add_argument("--seed", type=int, default=5)
    parser.add_argument("--iterations", type=int, default=200)
    parser.add_argument("--out", default=str(WEIGHTS))
    parser.add_argument("--init", default=None, help="start from these weights")
    parser.add_argument("--psf-size", type=int, nargs=3, default=list(PSF_SIZE), metavar=("Z", "Y", "X"),
                        help="support of the Gaussian PSF, sigma (3, 1.5, 1.5)")
    parser.add_argument("--peak", type=float, nargs=2, default=[PEAK, PEAK], metavar=("MIN", "MAX"),
                        help="photon peak drawn uniformly from [MIN, MAX] per pair")
    parser.add_argument("--blobs-per-mvox", type=float, default=BLOBS_PER_MEGAVOXEL)
    args = parser.parse_args(argv)
    net = warm_start_net(seed=args.seed)
    if args.init:
        net.load(args.init)
        net.step_count = 0
    t0 = time.perf_counter()
    shape = (args.edge,) * 3
    ys, xs = make_pairs(args.pairs, shape, args.iterations, args.seed, psf=psf_kernel(PSF_SIGMA, tuple(args.psf_size)),
                        peak_range=tuple(args.peak), blobs_per_megavoxel=args.blobs_per_mvox)
    print(f"[train] {args.pairs} pairs of {shape} in {time.perf_counter() - t0:.1f} s", flush=True)
    t0 = time.perf_counter()
    train(net, ys, xs, args.steps, lr=args.lr, log_every=max(1, args.steps // 40),
          log=lambda m: print(m, flush=True))
    print(f"[train] {args.steps} steps in {time.perf_counter() - t0:.1f} s", flush=True)
    net.save(args.out)
    print(f"[train] saved {args.out}", flush=True)


if __name__ == "__main__":
    main()
