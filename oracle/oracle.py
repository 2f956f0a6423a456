"""TEST INFRASTRUCTURE — CPU oracle of the IANE convolution / deconvolution path.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg import this module, and only as the checker (or the timed CPU baseline).
The product (``iane_amd``) never does.

Pieces and what pins each (see tests/test_oracle.py):
  * ``convolve``      — liboracle.so, the C restatement of the reference CPU
    convolve (include/iane/algorithms/convolution.hpp:23-117, locator/image_locator.hpp:40-56,
    algorithms/for_each.hpp:75-134). Pinned by the reference's known-answer
    cases, brute force, closed forms and by the reference itself (oracle/_ref).
  * ``convolve_ref``  — oracle/_ref/libiane_ref.so, the reference's own headers
    compiled here (present only where it was built).
  * ``fft*``          — FFT conventions of algorithms/fft/fft.hpp:196-231 and
    rfft.hpp:107-161 (negative exponent forward, unnormalised inverse, packed
    half spectrum along axis 0) restated over numpy.fft in float64; pinned by a
    brute-force DFT and the reference's FFT known answers.
  * ``rl`` / ``wiener`` — the loops of examples/rl_deconvolution/main.cu:51-81
    and examples/wiener/main.cu:21-83 in float64 (numpy); pinned against the
    reference loop run through oracle/_ref and the EM / identity properties.
  * ``sequence``      — the sequential fold of algorithms/detail/sequence.hpp:73-150
    (reduce / scans, whole or along an axis) with sum, product, min or max in
    the accumulator type; pinned by the closed forms of the reference's
    test/unit/prepared/sequence_cases.hpp (tests/test_oracle_sequence.py).

Array convention: numpy arrays in C order; IANE axis 0 is the LAST array axis.
"""
from __future__ import annotations

import ctypes
import os
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ORACLE_LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libiane_ref.so"

CLAMP, ZERO = 0, 1
_TYPES = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.int32): 2,
          np.dtype(np.float32): 3, np.dtype(np.float64): 4}


class _oview(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("rank", ctypes.c_int), ("dtype", ctypes.c_int),
                ("extent", ctypes.c_int * 4), ("stride", ctypes.c_int64 * 4)]


class _rview(ctypes.Structure):
    _fields_ = [("ptr", ctypes.c_void_p), ("rank", ctypes.c_int), ("dtype", ctypes.c_int),
                ("extent", ctypes.c_int * 4), ("stride", ctypes.c_int * 4)]


def _describe(a: np.ndarray, cls):
    v = cls()
    v.ptr = a.ctypes.data
    v.rank = a.ndim
    v.dtype = _TYPES[a.dtype]
    for axis in range(a.ndim):
        dim = a.ndim - 1 - axis
        v.extent[axis] = a.shape[dim]
        v.stride[axis] = a.strides[dim]
    return v


_olib = None
_rlib = None


def _oracle():
    global _olib
    if _olib is None:
        if not ORACLE_LIB.exists():
            raise ImportError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        lib = ctypes.CDLL(str(ORACLE_LIB))
        lib.oracle_convolve.argtypes = [ctypes.POINTER(_oview)] * 3 + [ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t]
        lib.oracle_convolve.restype = ctypes.c_int
        _olib = lib
    return _olib


def have_reference() -> bool:
    return REF_LIB.exists()


def _reference():
    global _rlib
    if _rlib is None:
        if not REF_LIB.exists():
            raise ImportError(f"{REF_LIB} missing (reference headers were not available to build it)")
        lib = ctypes.CDLL(str(REF_LIB))
        lib.ref_convolve.argtypes = [ctypes.POINTER(_rview)] * 3 + [ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t]
        lib.ref_convolve.restype = ctypes.c_int
        lib.ref_rl.argtypes = [ctypes.POINTER(_rview)] * 4 + [ctypes.c_int, ctypes.c_float, ctypes.c_size_t]
        lib.ref_rl.restype = ctypes.c_int
        lib.ref_last_error.restype = ctypes.c_char_p
        _rlib = lib
    return _rlib


def convolve(src: np.ndarray, kernel: np.ndarray, out_dtype=np.float64, border: int = CLAMP,
             threads: int = 1, grain: int = 4096, out: np.ndarray | None = None) -> np.ndarray:
    """Reference convolve semantics; ``out_dtype=float64`` returns the unrounded FP64 sum."""
    if out is None:
        out = np.zeros(src.shape, dtype=out_dtype)
    s, k, d = _describe(src, _oview), _describe(kernel, _oview), _describe(out, _oview)
    status = _oracle().oracle_convolve(ctypes.byref(s), ctypes.byref(k), ctypes.byref(d), border, threads, grain)
    if status:
        raise ValueError("Convolution requires positive odd kernel dimensions and matching extents")
    return out


def convolve_ref(src: np.ndarray, kernel: np.ndarray, out_dtype=np.float32, border: int = CLAMP,
                 threads: int = 1, grain: int = 4096, out: np.ndarray | None = None) -> np.ndarray:
    """The reference's own iane::convolve (compiled from its headers)."""
    lib = _reference()
    if out is None:
        out = np.zeros(src.shape, dtype=out_dtype)
    s, k, d = _describe(src, _rview), _describe(kernel, _rview), _describe(out, _rview)
    status = lib.ref_convolve(ctypes.byref(s), ctypes.byref(k), ctypes.byref(d), border, threads, grain)
    if status == 1:
        raise ValueError(lib.ref_last_error().decode())
    if status:
        raise RuntimeError(lib.ref_last_error().decode())
    return out


def rl_ref(observed: np.ndarray, psf: np.ndarray, iterations: int, epsilon: float = 1e-6,
           threads: int = 1) -> np.ndarray:
    """examples/rl_deconvolution/main.cu:51-81 run through the reference's primitives (float32)."""
    lib = _reference()
    observed = np.ascontiguousarray(observed, dtype=np.float32)
    psf = np.ascontiguousarray(psf, dtype=np.float32)
    flipped = np.ascontiguousarray(psf.reshape(-1)[::-1].reshape(psf.shape))
    out = np.zeros_like(observed)
    views = [_describe(a, _rview) for a in (observed, psf, flipped, out)]
    status = lib.ref_rl(*[ctypes.byref(v) for v in views], iterations, epsilon, threads)
    if status:
        raise RuntimeError(lib.ref_last_error().decode())
    return out


# ---------------------------------------------------------------- FFT -------
def fft_forward(x: np.ndarray) -> np.ndarray:
    """C2C, negative exponent, over every axis (fft.hpp:228-229)."""
    return np.fft.fftn(np.asarray(x, dtype=np.complex128))


def fft_inverse(x: np.ndarray) -> np.ndarray:
    """Unnormalised inverse (fft.hpp:230-231, docs/prepared_algorithms.md:26-29)."""
    x = np.asarray(x, dtype=np.complex128)
    return np.fft.ifftn(x) * x.size


def rfft(x: np.ndarray) -> np.ndarray:
    """R2C packed half spectrum along IANE axis 0 (= last array axis), rfft_extent (layout.hpp:15-21)."""
    return np.fft.rfftn(np.asarray(x, dtype=np.float64))


def irfft(spectrum: np.ndarray, real_shape) -> np.ndarray:
    """Unnormalised C2R inverse of ``rfft``."""
    n = int(np.prod(real_shape))
    return np.fft.irfftn(np.asarray(spectrum, dtype=np.complex128), s=real_shape, axes=list(range(len(real_shape)))) * n


def dft_bruteforce(x: np.ndarray) -> np.ndarray:
    """Direct O(N^2) FP64 DFT with the negative exponent (real_fft_cases.hpp:44-57 style)."""
    x = np.asarray(x, dtype=np.complex128)
    out = x
    for axis in range(x.ndim):
        n = x.shape[axis]
        k = np.arange(n)
        w = np.exp(-2j * np.pi * np.outer(k, k) / n)
        out = np.moveaxis(np.tensordot(w, np.moveaxis(out, axis, 0), axes=(1, 0)), 0, axis)
    return out


# ---------------------------------------------------------- RL / Wiener -----
def _fft_workers() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def conv_zero_fft(x: np.ndarray, psf: np.ndarray) -> np.ndarray:
    """Zero-border convolution with an odd centred kernel, as a float64 FFT
    product on a padded domain (equal to ``convolve(..., border=ZERO)`` in exact arithmetic).
    scipy's pocketfft in float64, threaded over the host cores."""
    from scipy import fft as sfft
    shape = [n + k - 1 for n, k in zip(x.shape, psf.shape)]
    axes = list(range(x.ndim))
    w = _fft_workers()
    x = np.asarray(x, dtype=np.float64)
    psf = np.asarray(psf, dtype=np.float64)
    full = sfft.irfftn(sfft.rfftn(x, shape, axes, workers=w) * sfft.rfftn(psf, shape, axes, workers=w), shape, axes,
                       workers=w)
    sl = tuple(slice(k // 2, k // 2 + n) for n, k in zip(x.shape, psf.shape))
    return full[sl]


def poisson_ll(observed: np.ndarray, predicted: np.ndarray, epsilon: float = 1e-6) -> float:
    """Poisson log-likelihood up to the constant -sum(log y!): sum(y log max(p, eps) - p)."""
    y = np.asarray(observed, dtype=np.float64)
    p = np.asarray(predicted, dtype=np.float64)
    return float(np.sum(y * np.log(np.maximum(p, epsilon)) - p))


def rl(observed: np.ndarray, psf: np.ndarray, iterations: int, epsilon: float = 1e-6,
       init: np.ndarray | None = None, direct: bool = False):
    """FP64 restatement of examples/rl_deconvolution/main.cu:51-81.

    Returns ``(estimate, ll)`` with ``ll[i]`` the log-likelihood of iteration i's
    prediction H x_i. ``direct=True`` uses the C oracle convolve (small sizes)."""
    y = np.asarray(observed, dtype=np.float64)
    h = np.asarray(psf, dtype=np.float64)
    flipped = h[tuple(slice(None, None, -1) for _ in range(h.ndim))].copy()

    def conv(a, k):
        if direct:
            return convolve(np.ascontiguousarray(a), np.ascontiguousarray(k), np.float64, ZERO)
        return conv_zero_fft(a, k)

    sensitivity = conv(np.ones_like(y), flipped)
    x = np.where(y > epsilon, y, epsilon) if init is None else np.asarray(init, dtype=np.float64).copy()
    ll = []
    for _ in range(iterations):
        p = conv(x, h)
        ll.append(poisson_ll(y, p, epsilon))
        ratio = y / np.where(p > epsilon, p, epsilon)
        c = conv(ratio, flipped)
        factor = np.where(sensitivity > 0, c / np.where(sensitivity > epsilon, sensitivity, epsilon), 0.0)
        x = x * factor
    return x, np.array(ll)


def wiener(observed: np.ndarray, psf_padded: np.ndarray, noise: float) -> np.ndarray:
    """examples/wiener/main.cu:44-83 in float64: X = Y conj(H) / (|H|^2 + noise), x = Re(IFFT X) / N."""
    y = np.fft.fftn(np.asarray(observed, dtype=np.float64))
    h = np.fft.fftn(np.asarray(psf_padded, dtype=np.complex128))
    x = y * np.conj(h) / (np.abs(h) ** 2 + noise)
    return np.real(np.fft.ifftn(x))


def pad_psf_centered(psf: np.ndarray, shape) -> np.ndarray:
    """Place the PSF centre at the origin of an image-sized buffer (wiener/main.cu:109-120)."""
    out = np.zeros(shape, dtype=np.float64)
    idx = np.indices(psf.shape).reshape(psf.ndim, -1)
    target = tuple((idx[a] - psf.shape[a] // 2) % shape[a] for a in range(psf.ndim))
    out[target] = psf.reshape(-1)
    return out


# ------------------------------------------------------------ sequences -----
SEQ_INCLUSIVE, SEQ_EXCLUSIVE, SEQ_REDUCE = 0, 1, 2
_SEQ_UFUNC = {"sum": np.add, "product": np.multiply, "min": np.minimum, "max": np.maximum}


def sequence(src: np.ndarray, kind: int, axis: int, op: str, init, acc_dtype) -> np.ndarray:
    """The serial fold of algorithms/detail/sequence.hpp:73-150: every element is
    converted to the accumulator type, ``total = init`` and ``total = op(total, x)``
    in storage order along each line (axis -1: the whole array in C order, IANE
    axis a = array axis ndim-1-a). Integer arithmetic wraps like the accumulator
    type. Returns the scan (source shape) or the reduction (axis removed; shape
    (1,) for a whole or 1-D reduction)."""
    acc = np.dtype(acc_dtype)
    ufunc = _SEQ_UFUNC[op]
    x = np.asarray(src).astype(acc)
    start = np.asarray(init).astype(acc)
    if axis < 0:
        lines = x.reshape(1, -1)
    else:
        lines = np.moveaxis(x, x.ndim - 1 - axis, -1)
    shape = lines.shape
    flat = lines.reshape(int(np.prod(shape[:-1], dtype=np.int64)), shape[-1])
    padded = np.concatenate([np.full((flat.shape[0], 1), start, dtype=acc), flat], axis=1)
    with np.errstate(over="ignore"):
        folded = ufunc.accumulate(padded, axis=1, dtype=acc)
    if kind == SEQ_REDUCE:
        totals = folded[:, -1]
        if axis < 0 or x.ndim == 1:
            return totals.reshape(1)
        return totals.reshape(shape[:-1])
    out = (folded[:, 1:] if kind == SEQ_INCLUSIVE else folded[:, :-1]).reshape(shape)
    if axis < 0:
        return out.reshape(x.shape)
    return np.moveaxis(out, -1, x.ndim - 1 - axis)

