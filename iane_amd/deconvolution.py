"""Richardson–Lucy and Wiener deconvolution on device images.

``rl_plan`` / ``prepare_rl`` mirror prepare_rl of
examples/rl_deconvolution/main.cu:51-81: the prepared state owns the PSF, its
flip (main.cu:106-107) or spectrum, and the sensitivity H^T 1 (main.cu:62-63);
running it performs ``iterations`` RL steps from max(y, eps) (main.cu:68) or
from a supplied warm start. ``wiener`` mirrors prepare_wiener of
examples/wiener/main.cu:44-83 on the package's FFT.
All arithmetic runs in libiane_hip.so.
"""
from __future__ import annotations

import ctypes

from . import _lib, hip_execution, hip_policy, precision as precision_t, view_desc
from ._lib import invalid_argument

DIRECT, FFT = 0, 1


def _stream(policy):
    if policy is None:
        policy = hip_policy()
    if isinstance(policy, hip_execution):
        policy = policy.policy()
    return policy.handle()


class rl_plan:
    """Prepared Richardson–Lucy for one image shape and PSF (rank 2 or 3, FP32)."""

    def __init__(self, shape, psf, method: int = FFT, precision: int = precision_t.fp32,
                 epsilon: float = 1e-6, policy=None):
        import torch
        shape = tuple(int(s) for s in shape)
        if len(shape) not in (2, 3):
            raise _lib.unsupported("RL plans support rank 2 and 3 images")
        self.shape = shape
        self.method = int(method)
        self.precision = int(precision)
        self.epsilon = float(epsilon)
        self.stream = _stream(policy)
        extent = (ctypes.c_int32 * 3)(*([s for s in reversed(shape)] + [1] * (3 - len(shape))))
        handle = ctypes.c_void_p()
        _lib.call("iane_hip_rl_plan_create", ctypes.byref(handle), len(shape), extent, view_desc(psf),
                  self.method, self.precision, ctypes.c_float(self.epsilon), self.stream)
        self._handle = handle
        nbytes = ctypes.c_size_t()
        _lib.call("iane_hip_rl_plan_workspace_bytes", handle, ctypes.byref(nbytes))
        self.workspace_bytes = nbytes.value
        self._workspace = torch.empty(max(self.workspace_bytes, 1), dtype=torch.uint8, device="cuda")

    def run(self, observed, estimate, iterations: int, *, init: bool = True, ll_history=None,
            workspace=None) -> hip_execution:
        """``iterations`` RL steps; ``ll_history`` (float64 device tensor of
        length >= iterations) receives the log-likelihood of each prediction."""
        import torch
        for t, name in ((observed, "observed"), (estimate, "estimate")):
            if t.device.type != "cuda" or t.dtype != torch.float32 or not t.is_contiguous():
                raise invalid_argument(f"{name} must be a packed float32 device image")
            if tuple(t.shape) != self.shape:
                raise invalid_argument("Image dimensions must match the prepared shape")
        if iterations < 0:
            raise invalid_argument("Iterations and finite epsilon must be positive")
        ll_ptr = 0
        if ll_history is not None:
            if ll_history.dtype != torch.float64 or ll_history.numel() < iterations or ll_history.device.type != "cuda":
                raise invalid_argument("ll_history must be a float64 device tensor with one entry per iteration")
            ll_ptr = ll_history.data_ptr()
        ws = workspace if workspace is not None else self._workspace
        if ws.device.type != "cuda" or ws.numel() * ws.element_size() < self.workspace_bytes:
            raise _lib.length_error("Prepared operation exceeded its scratch reservation")
        _lib.call("iane_hip_rl_run", self._handle, observed.data_ptr(), estimate.data_ptr(), int(iterations),
                  1 if init else 0, ws.data_ptr(), ll_ptr, self.stream)
        return hip_execution(self.stream)

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle:
            try:
                _lib.load().iane_hip_stream_sync(self.stream)
                _lib.load().iane_hip_rl_plan_destroy(handle)
            except Exception:
                pass
            self._handle = None


class rl_batch:
    """Many RL problems of uneven shapes with their own PSFs, advanced together:
    one launch per pass over every volume, the run captured once into a hipGraph
    and replayed. Per volume identical to ``rl_plan(method=FFT)``."""

    def __init__(self, shapes, psfs, epsilon: float = 1e-6, policy=None):
        import torch
        shapes = [tuple(int(s) for s in sh) for sh in shapes]
        if not shapes or len(shapes) != len(psfs):
            raise invalid_argument("one PSF per volume is required")
        rank = len(shapes[0])
        if any(len(s) != rank for s in shapes):
            raise invalid_argument("all volumes of a batch have the same rank")
        self.shapes = shapes
        self.stream = _stream(policy)
        extents = (ctypes.c_int32 * (rank * len(shapes)))(*[e for s in shapes for e in reversed(s)])
        descs = (_lib.iane_view_desc * len(psfs))(*[view_desc(p) for p in psfs])
        handle = ctypes.c_void_p()
        _lib.call("iane_hip_rl_batch_create", ctypes.byref(handle), len(shapes), rank, extents, descs,
                  ctypes.c_float(epsilon), self.stream)
        self._handle = handle
        nbytes = ctypes.c_size_t()
        _lib.call("iane_hip_rl_batch_workspace_bytes", handle, ctypes.byref(nbytes))
        self.workspace_bytes = nbytes.value
        self._workspace = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device="cuda")

    def run(self, observed, estimates, iterations: int, *, init: bool = True, ll_history=None,
            use_graph: bool | None = None) -> hip_execution:
        """``ll_history``: float64 device tensor of shape (count, iterations) or None.
        ``use_graph`` (default: whenever the batch runs on a created stream) replays
        the iteration sequence as one hipGraph; graph capture needs a created stream."""
        if use_graph is None:
            use_graph = self.stream != 0
        import torch
        n = len(self.shapes)
        if len(observed) != n or len(estimates) != n:
            raise invalid_argument("one observed and one estimate image per volume")
        for o, e, s in zip(observed, estimates, self.shapes):
            for t in (o, e):
                if tuple(t.shape) != s or t.dtype != torch.float32 or not t.is_contiguous() or t.device.type != "cuda":
                    raise invalid_argument("volumes must be packed float32 device images of the prepared shapes")
        obs = (ctypes.c_void_p * n)(*[o.data_ptr() for o in observed])
        est = (ctypes.c_void_p * n)(*[e.data_ptr() for e in estimates])
        ll_ptr = 0
        if ll_history is not None:
            if ll_history.dtype != torch.float64 or tuple(ll_history.shape) != (n, iterations) \
                    or not ll_history.is_contiguous() or ll_history.device.type != "cuda":
                raise invalid_argument("ll_history must be a contiguous float64 (count, iterations) tensor")
            ll_ptr = ll_history.data_ptr()
        self._keep = (obs, est)
        _lib.call("iane_hip_rl_batch_run", self._handle, obs, est, int(iterations), 1 if init else 0,
                  self._workspace.data_ptr(), ll_ptr, 1 if use_graph else 0, self.stream)
        return hip_execution(self.stream)

    def __del__(self):
        handle = getattr(self, "_handle", None)
        if handle:
            try:
                _lib.load().iane_hip_stream_sync(self.stream)
                _lib.load().iane_hip_rl_batch_destroy(handle)
            except Exception:
                pass
            self._handle = None


def richardson_lucy(observed, psf, iterations: int = 20, epsilon: float = 1e-6, *, method: int = FFT,
                    precision: int = precision_t.fp32, warm_start=None, return_ll: bool = False, policy=None):
    """One-shot RL: returns the estimate (and the log-likelihood history)."""
    import torch
    plan = rl_plan(observed.shape, psf, method, precision, epsilon, policy)
    estimate = warm_start.clone() if warm_start is not None else torch.empty_like(observed)
    # entries 0..iterations-1 are all written by the run (on the plan's stream)
    ll = torch.empty(max(iterations, 1), dtype=torch.float64, device="cuda") if return_ll else None
    plan.run(observed, estimate, iterations, init=warm_start is None, ll_history=ll).wait()
    return (estimate, ll[:iterations]) if return_ll else estimate


def _check_pointwise(*images):
    """Pointwise steps take packed float32 device images of one element count."""
    import torch
    count = images[0][0].numel()
    for t, name in images:
        if t.device.type != "cuda" or t.dtype != torch.float32 or not t.is_contiguous():
            raise invalid_argument(f"{name} must be a packed float32 device image")
        if t.numel() != count:
            raise invalid_argument("pointwise images require matching element counts")


def rl_initial(observed, estimate, epsilon: float = 1e-6, policy=None) -> hip_execution:
    """estimate = max(observed, eps) (main.cu:12-19)."""
    _check_pointwise((observed, "observed"), (estimate, "estimate"))
    s = _stream(policy)
    _lib.call("iane_hip_rl_initial", observed.data_ptr(), estimate.data_ptr(), observed.numel(),
              ctypes.c_float(epsilon), s)
    return hip_execution(s)


def rl_ratio(observed, predicted, ratio, epsilon: float = 1e-6, ll_out=None, policy=None,
             workspace=None) -> hip_execution:
    """ratio = observed / max(predicted, eps) (main.cu:22-29); ``ll_out`` (a float64
    device tensor) gets the log-likelihood, summed through ``workspace``."""
    import torch
    _check_pointwise((observed, "observed"), (predicted, "predicted"), (ratio, "ratio"))
    s = _stream(policy)
    ll_ptr, ws_ptr, ws_bytes = 0, 0, 0
    if ll_out is not None:
        if ll_out.device.type != "cuda" or ll_out.dtype != torch.float64 or ll_out.numel() < 1:
            raise invalid_argument("ll_out must be a float64 device tensor with at least one element")
        ll_ptr = ll_out.data_ptr()
        nbytes = ctypes.c_size_t()
        _lib.call("iane_hip_rl_ratio_workspace_bytes", observed.numel(), ctypes.byref(nbytes))
        if workspace is None:
            workspace = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device="cuda")
        ws_bytes = workspace.numel() * workspace.element_size()
        if workspace.device.type != "cuda" or ws_bytes < nbytes.value:
            raise _lib.length_error("Prepared operation exceeded its scratch reservation")
        ws_ptr = workspace.data_ptr()
    _lib.call("iane_hip_rl_ratio", observed.data_ptr(), predicted.data_ptr(), ratio.data_ptr(), observed.numel(),
              ctypes.c_float(epsilon), ll_ptr, ws_ptr, ws_bytes, s)
    return hip_execution(s)


def rl_update(estimate, correction, sensitivity, epsilon: float = 1e-6, policy=None) -> hip_execution:
    """estimate *= (s > 0 ? c / max(s, eps) : 0) (main.cu:32-49)."""
    _check_pointwise((estimate, "estimate"), (correction, "correction"), (sensitivity, "sensitivity"))
    s = _stream(policy)
    _lib.call("iane_hip_rl_update", estimate.data_ptr(), correction.data_ptr(), sensitivity.data_ptr(),
              estimate.numel(), ctypes.c_float(epsilon), s)
    return hip_execution(s)


def wiener_apply(spectrum, response, noise: float, policy=None) -> hip_execution:
    """spectrum *= conj(H) / (|H|^2 + noise) in place (wiener/main.cu:21-32)."""
    s = _stream(policy)
    _lib.call("iane_hip_wiener_apply", spectrum.data_ptr(), response.data_ptr(), spectrum.numel(),
              ctypes.c_float(noise), s)
    return hip_execution(s)


def wiener(observed, psf_padded, noise: float = 0.002, policy=None):
    """prepare_wiener + closure (wiener/main.cu:44-83) for a real float image:
    FFT(y) * conj(H) / (|H|^2 + noise), inverse FFT, real part / N.
    ``psf_padded``: complex64 image-size PSF with its centre at the origin (main.cu:109-120)."""
    import torch
    from .prepared import execution_context, fft_forward, fft_inverse
    if not (noise > 0) or noise != noise or noise == float("inf"):
        raise invalid_argument("Noise must be finite and positive")
    stream = _stream(policy)
    ctx = execution_context(hip_policy(stream))
    spatial = torch.empty(observed.shape, dtype=torch.complex64, device="cuda")
    spectrum = torch.empty_like(spatial)
    response = torch.empty_like(spatial)
    out = torch.empty(observed.shape, dtype=torch.float32, device="cuda")
    fwd = ctx.prepare(fft_forward, spatial, spectrum)
    inv = ctx.prepare(fft_inverse, spectrum, spatial)
    ctx.run(fwd, psf_padded, response)
    _lib.call("iane_hip_real_to_complex", observed.data_ptr(), spatial.data_ptr(), observed.numel(), stream)
    ctx.run(fwd, spatial, spectrum)
    wiener_apply(spectrum, response, noise, ctx.policy)
    ctx.run(inv, spectrum, spatial)
    scale = 1.0 / float(observed.numel())  # wiener/main.cu:72 (computed in double, stored as float)
    _lib.call("iane_hip_complex_real_scale", spatial.data_ptr(), out.data_ptr(), observed.numel(),
              ctypes.c_float(scale), stream)
    hip_execution(stream).wait()
    return out
