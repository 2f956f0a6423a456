"""Texture images as convolution sources (include/iane/resource/cuda/texture_image.hpp,
include/iane/view/texture_image_view.hpp; test_convolution_cuda.cu:14-163 of the reference).

REFERENCE precision through a texture view is bit-identical to the oracle on the
cropped image (the crop is an independent image for the border rule), for every
texel type; FP32 precision is bit-identical to the same call on a memory view.
"""
from __future__ import annotations

import numpy as np
import pytest

import iane_amd
from iane_amd import border_mode, convolution_options, precision
from oracle import oracle
from tests import cases


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def same_bits(a, b):
    np.testing.assert_array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


# ------------------------------------------------------------ host-only -----
def test_texture_validation_without_gpu():
    import torch
    with pytest.raises(iane_amd.invalid_argument):
        iane_amd.texture_image((2, -1))
    with pytest.raises(iane_amd.unsupported):
        iane_amd.texture_image((2, 2, 2, 2))
    with pytest.raises(iane_amd.unsupported):
        iane_amd.texture_image((4,), torch.float64)
    empty = iane_amd.texture_image((3, 0))   # empty_texture_and_move: no allocation
    assert empty.array() is None and empty.dimensions() == (0, 3)
    empty.upload(torch.empty((3, 0)))        # a matching empty upload is a no-op
    with pytest.raises(iane_amd.invalid_argument):
        empty.upload(torch.empty((2, 0)))


# --------------------------------------------------------------- device -----
@pytest.mark.gpu
def test_upload_from_strided_host_and_device_views():
    import torch
    host = (torch.arange(12, dtype=torch.int32) + 101).reshape(3, 4)
    texture = iane_amd.texture_image((2, 2), torch.int32)
    with pytest.raises(iane_amd.invalid_argument):
        iane_amd.copy(host, texture)                      # extent mismatch
    out = torch.empty((2, 2), dtype=torch.int32, device="cuda")

    def verify(offset):
        iane_amd.copy(texture.const_view(), out).wait()
        assert out.cpu().flatten().tolist() == [106 + offset, 107 + offset, 110 + offset, 111 + offset]

    iane_amd.copy(host[1:3, 1:3], texture).wait()         # strided host crop
    verify(0)
    linear = (host + 100).cuda()
    iane_amd.copy(linear[1:3, 1:3], texture).wait()       # strided device crop
    verify(100)
    with pytest.raises(iane_amd.invalid_argument):
        iane_amd.copy(linear.t()[1:3, 1:3].contiguous().t(), texture)  # columns, not rows, contiguous


@pytest.mark.gpu
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_texture_convolution_matches_cpu(border):
    """test_convolution_cuda.cu:53-86: the crop of a texture convolves exactly like the host crop."""
    import torch
    source, crop, kernel = cases.texture_case()
    texture = iane_amd.texture_image(source.shape, torch.float32)
    iane_amd.copy(torch.from_numpy(source), texture).wait()
    out = torch.full((3, 3), 77.0, device="cuda")
    iane_amd.convolve(texture.const_view().subview((1, 1), (3, 3)), dev(kernel), out,
                      convolution_options(border_mode(border))).wait()
    same_bits(out.cpu().numpy(), oracle.convolve(np.ascontiguousarray(crop), kernel, np.float32, border))


@pytest.mark.gpu
def test_1d_upload_and_nested_crop():
    import torch
    source = torch.arange(100, 106, dtype=torch.int32)
    texture = iane_amd.texture_image((4,), torch.int32)
    iane_amd.copy(source[1:5], texture).wait()
    out = torch.empty(2, dtype=torch.int32, device="cuda")
    iane_amd.copy(texture.const_view().subview((1,), (3,)).subview((1,), (2,)), out).wait()
    assert out.cpu().tolist() == [103, 104]


@pytest.mark.gpu
def test_3d_upload_slice_stride_not_a_whole_number_of_rows():
    import torch
    padded = torch.zeros(33, dtype=torch.int32)
    host = padded.as_strided((3, 2, 3), (11, 4, 1))  # slice 44 B, pitch 16 B
    for z in range(3):
        for y in range(2):
            for x in range(3):
                host[z, y, x] = 100 * z + 10 * y + x
    texture = iane_amd.texture_image((3, 2, 3), torch.int32)
    iane_amd.copy(host, texture).wait()
    out = torch.empty((2, 2, 2), dtype=torch.int32, device="cuda")
    iane_amd.copy(texture.const_view().subview((1, 0, 1), (2, 2, 2)), out).wait()
    got = out.cpu()
    for z in range(2):
        for y in range(2):
            for x in range(2):
                assert int(got[z, y, x]) == 100 * (z + 1) + 10 * y + x + 1
    storage = torch.full((4, 4, 5), -1, dtype=torch.int32)
    storage[1:4, 1:3, 1:4] = host + 1000
    linear = storage.cuda()
    iane_amd.copy(linear[1:4, 1:3, 1:4], texture).wait()
    iane_amd.copy(texture.const_view().subview((1, 0, 1), (2, 2, 2)), out).wait()
    assert int(out.cpu()[0, 0, 0]) == 1101


@pytest.mark.gpu
@pytest.mark.parametrize("tdt,kdt,ddt", [(np.uint8, np.float32, np.uint8), (np.uint16, np.int32, np.int32),
                                         (np.int32, np.int32, np.int32), (np.float32, np.float32, np.float32)])
@pytest.mark.parametrize("border", [cases.CLAMP, cases.ZERO])
def test_cropped_volume_bit_exact_every_texel_type(tdt, kdt, ddt, border):
    import torch
    rng = np.random.default_rng(7)
    shape = (20, 37, 70)
    if np.issubdtype(tdt, np.integer):
        vol = rng.integers(0, 200, shape).astype(tdt)
    else:
        vol = rng.standard_normal(shape).astype(tdt)
    ker = (rng.integers(-3, 4, (3, 5, 7)) if np.issubdtype(kdt, np.integer) else rng.standard_normal((3, 5, 7))).astype(kdt)
    texture = iane_amd.texture_image(shape, torch.from_numpy(vol).dtype)
    iane_amd.copy(torch.from_numpy(vol), texture).wait()
    corner, size = (3, 5, 9), (14, 25, 50)
    crop = vol[3:17, 5:30, 9:59]
    out = dev(np.full(size, 77, ddt))
    iane_amd.convolve(texture.const_view().subview(corner, size), dev(ker), out,
                      convolution_options(border_mode(border))).wait()
    same_bits(out.cpu().numpy(), oracle.convolve(np.ascontiguousarray(crop), ker, ddt, border))


@pytest.mark.gpu
def test_fp32_texture_equals_memory_view():
    import torch
    rng = np.random.default_rng(8)
    vol = rng.random((16, 64, 96)).astype(np.float32)
    ker = rng.random((5, 5, 5)).astype(np.float32)
    texture = iane_amd.texture_image(vol.shape, torch.float32)
    iane_amd.copy(dev(vol), texture).wait()
    opts = convolution_options(border_mode.zero, precision.fp32)
    a = torch.empty((10, 40, 80), device="cuda")
    b = torch.empty_like(a)
    iane_amd.convolve(texture.const_view().subview((2, 10, 8), (10, 40, 80)), dev(ker), a, opts).wait()
    iane_amd.convolve(dev(vol)[2:12, 10:50, 8:88], dev(ker), b, opts).wait()
    same_bits(a.cpu().numpy(), b.cpu().numpy())
