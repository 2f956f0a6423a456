"""Split-precision MFMA convolution (precision.split_fp16 / split_bf16) against the FP64 oracle.

Error metric: e = max|gpu - oracle64| / max|oracle64| (and the normwise
relative L2 error, printed). Bounds, from the operand split:
  split_fp16: hi + lo, 11 + 11 significant bits per operand, FP32 accumulation -> TOL_FP16 = 1e-5
              (the same bound as the FP32 direct path)
  split_bf16: hi + mid + lo, 8 + 8 + 8 significant bits per operand, the six
              products down to 2^-16 of the leading one                    -> TOL_BF16 = 1e-5
The path needs finite inputs (the operands are scaled by powers of two
derived from max|src| and max|kernel|).
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import iane_amd
from iane_amd import border_mode, convolution_options, precision
from oracle import oracle
from tests import cases

pytestmark = pytest.mark.gpu
TOL_FP16 = 1e-5
TOL_BF16 = 1e-5
TOL = {precision.split_fp16: TOL_FP16, precision.split_bf16: TOL_BF16}


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def run(src, ker, border, prec, ddt=np.float32):
    d = dev(np.full(src.shape, 77, ddt))
    s, k = dev(src), dev(ker)
    iane_amd.convolve(s, k, d, convolution_options(border_mode(border), prec)).wait()
    return d.cpu().numpy()


def errors(got, ref):
    got = got.astype(np.float64)
    emax = float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))
    el2 = float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))
    return emax, el2


@pytest.mark.parametrize("prec", [precision.split_fp16, precision.split_bf16])
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_shifted_delta_is_a_shift(prec, border):
    """A single unit tap at kernel offset d moves the image by d - centre
    (closed form), border rule included; exercises orientation and padding."""
    rng = np.random.default_rng(1)
    src = rng.random((9, 21, 270)).astype(np.float32)
    ker = np.zeros((5, 7, 15), np.float32)
    ker[1, 5, 2] = 1.0
    got = run(src, ker, border, prec)
    ref = oracle.convolve(src, ker, np.float64, border)
    emax, _ = errors(got, ref)
    # one product = the input's split representation: 22 bits (FP16) or 24 bits (BF16)
    assert emax <= (2 ** -20 if prec == precision.split_fp16 else 2 ** -22)


@pytest.mark.parametrize("prec", [precision.split_fp16, precision.split_bf16])
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
@pytest.mark.parametrize("kshape", [(3, 3, 3), (9, 7, 5), (15, 15, 15), (3, 17, 9), (7, 9, 31), (1, 1, 1)])
def test_random_volumes(prec, border, kshape):
    rng = np.random.default_rng(sum(kshape))
    src = (rng.random((10, 37, 300)) + 0.1 * rng.standard_normal((10, 37, 300))).astype(np.float32)
    ker = rng.standard_normal(kshape).astype(np.float32)
    ker /= np.float32(np.abs(ker).sum())
    got = run(src, ker, border, prec)
    ref = oracle.convolve(src, ker, np.float64, border, threads=8)
    emax, el2 = errors(got, ref)
    print(f"{prec.name} {kshape} border={border}: max-rel {emax:.3e} l2-rel {el2:.3e}")
    assert emax <= TOL[prec]


@pytest.mark.parametrize("prec", [precision.split_fp16, precision.split_bf16])
def test_dense15_blob_volume(prec):
    """The dense random 15^3 kernel (normal, normalised to sum 1) on a blob volume."""
    vol = __import__("iane_amd.synth", fromlist=["x"]).volume_blobs(40, seed=2, blobs=8)
    ref = oracle.convolve(vol["volume"], vol["dense"], np.float64, cases.ZERO, threads=8)
    got = run(vol["volume"], vol["dense"], cases.ZERO, prec)
    emax, el2 = errors(got, ref)
    print(f"{prec.name} dense15: max-rel {emax:.3e} l2-rel {el2:.3e}")
    assert emax <= TOL[prec]


@pytest.mark.parametrize("prec,tol", [(precision.fp32, 1e-5), (precision.split_fp16, TOL_FP16),
                                      (precision.split_bf16, TOL_BF16)], ids=["fp32", "split_fp16", "split_bf16"])
def test_dense15_full_size_vs_fft_oracle(prec, tol):
    """The dense random 15^3 kernel (normal, normalised to sum 1) on the 256^3
    blob volume, zero border, against the FP64 zero-padded FFT convolution of
    the oracle (itself pinned to the direct FP64 convolve)."""
    from iane_amd import synth
    data = synth.volume_blobs(256, seed=2)
    got = run(data["volume"], data["dense"], cases.ZERO, prec)
    ref = oracle.conv_zero_fft(data["volume"].astype(np.float64), data["dense"].astype(np.float64))
    emax, el2 = errors(got, ref)
    print(f"dense 15^3 at 256^3, precision code {int(prec)}: max {emax:.2e}, normwise {el2:.2e}")
    assert emax <= tol


@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_two_dimensional_image(border):
    rng = np.random.default_rng(3)
    src = rng.random((130, 520)).astype(np.float32)
    ker = rng.random((7, 11)).astype(np.float32)
    ker /= ker.sum()
    emax, _ = errors(run(src, ker, border, precision.split_fp16), oracle.convolve(src, ker, np.float64, border))
    assert emax <= TOL_FP16


def test_strided_views_and_double_destination():
    import torch
    from iane_amd import _lib
    rng = np.random.default_rng(4)
    base = rng.random((12, 40, 90)).astype(np.float32)
    ker = rng.random((5, 5, 5)).astype(np.float32)
    ker /= ker.sum()
    dbase, dker = dev(base), dev(ker)
    view = dbase[::2, 3:35, :]                     # every other plane, cropped rows
    desc = iane_amd.view_desc(view)                # ... and x reversed (negative IANE axis-0 stride)
    desc.ptr = view.data_ptr() + (base.shape[2] - 1) * desc.stride_bytes[0]
    desc.stride_bytes[0] = -desc.stride_bytes[0]
    src = base[::2, 3:35, ::-1]
    out = torch.full((12, 32, 90), 77.0, dtype=torch.float64, device="cuda")
    dst_t = out[::2]                                # strided double destination
    nbytes = ctypes.c_size_t()
    _lib.call("iane_hip_convolve_workspace_bytes", desc, iane_amd.view_desc(dker), iane_amd.view_desc(dst_t),
              cases.CLAMP, int(precision.split_fp16), ctypes.byref(nbytes))
    assert nbytes.value > 0
    ws = torch.empty(nbytes.value, dtype=torch.uint8, device="cuda")
    # too small a workspace is refused before anything runs
    with pytest.raises(_lib.length_error):
        _lib.call("iane_hip_convolve", desc, iane_amd.view_desc(dker), iane_amd.view_desc(dst_t), cases.CLAMP,
                  int(precision.split_fp16), ws.data_ptr(), nbytes.value - 1, 0)
    _lib.call("iane_hip_convolve", desc, iane_amd.view_desc(dker), iane_amd.view_desc(dst_t), cases.CLAMP,
              int(precision.split_fp16), ws.data_ptr(), nbytes.value, 0)
    iane_amd.hip_execution(0).wait()
    ref = oracle.convolve(np.ascontiguousarray(src), ker, np.float64, cases.CLAMP)
    emax, _ = errors(dst_t.cpu().numpy(), ref)
    assert emax <= TOL_FP16
    assert np.all(out.cpu().numpy()[1::2] == 77.0)


def test_wide_kernel_falls_back_to_fp32():
    """Kernels wider than 49 along x are not tiled by the MFMA path; the call
    still succeeds (FP32 direct kernel) within the FP32 bound."""
    rng = np.random.default_rng(5)
    src = rng.random((3, 8, 120)).astype(np.float32)
    ker = rng.random((1, 3, 51)).astype(np.float32)
    ker /= ker.sum()
    emax, _ = errors(run(src, ker, cases.ZERO, precision.split_fp16), oracle.convolve(src, ker, np.float64, cases.ZERO))
    assert emax <= 1e-5


def test_zero_image_and_zero_kernel():
    src = np.zeros((4, 20, 40), np.float32)
    ker = np.ones((3, 3, 3), np.float32)
    assert np.all(run(src, ker, cases.ZERO, precision.split_fp16) == 0.0)
    src = np.ones((4, 20, 40), np.float32)
    assert np.all(run(src, np.zeros((3, 3, 3), np.float32), cases.CLAMP, precision.split_fp16) == 0.0)
