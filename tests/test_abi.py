"""The C-ABI library loads without a GPU and exports every symbol include/iane_hip.h declares."""
from __future__ import annotations

import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "iane_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(iane_hip_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "iane_hip_convolve" in names and "iane_hip_last_error" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    import iane_amd
    lib = iane_amd.library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_version_and_error_string_without_gpu():
    import iane_amd
    lib = iane_amd.library()
    assert lib.iane_hip_version() >= 1
    assert isinstance(lib.iane_hip_last_error(), bytes)


def test_invalid_descriptor_reports_invalid_argument():
    """Argument validation happens on the host before any HIP call."""
    import ctypes
    from iane_amd import _lib
    lib = _lib.load()
    status = lib.iane_hip_convolve(None, None, None, 0, 0, None, 0, None)
    assert status == _lib.ERR_INVALID_ARGUMENT
    assert b"null" in lib.iane_hip_last_error()
    d = _lib.iane_view_desc()
    d.rank = 2
    d.dtype = _lib.F32
    d.extent[0] = d.extent[1] = 4
    k = _lib.iane_view_desc()
    k.rank = 2
    k.dtype = _lib.F32
    k.extent[0], k.extent[1] = 2, 3  # even kernel extent
    status = lib.iane_hip_convolve(ctypes.byref(d), ctypes.byref(k), ctypes.byref(d), 0, 0, None, 0, None)
    assert status == _lib.ERR_INVALID_ARGUMENT
    assert b"odd" in lib.iane_hip_last_error()
    with pytest.raises(_lib.invalid_argument):
        _lib.check(status)


def test_separable_precision_contract_without_gpu():
    """The separable entry point takes REFERENCE, FP32 and SPLIT_FP16 (the split
    engine's fused path needs no workspace) and refuses SPLIT_BF16; the query runs
    on the host only."""
    import ctypes
    from iane_amd import _lib
    lib = _lib.load()

    def desc(extents, dtype=_lib.F32):
        d = _lib.iane_view_desc()
        d.ptr = 256  # never dereferenced by the query
        d.rank = len(extents)
        d.dtype = dtype
        stride = 4
        for i, e in enumerate(extents):
            d.extent[i] = e
            d.stride_bytes[i] = stride
            stride *= e
        return d

    src = desc((32, 20, 12))
    taps = (_lib.iane_view_desc * 3)(desc((5,)), desc((5,)), desc((5,)))
    size = ctypes.c_size_t(77)
    for prec, fused in ((0, False), (1, True), (3, True)):
        status = lib.iane_hip_convolve_separable_workspace_bytes(ctypes.byref(src), taps, 3, ctypes.byref(src), 1,
                                                                 prec, ctypes.byref(size))
        assert status == _lib.OK, lib.iane_hip_last_error()
        # FP32 / SPLIT_FP16 on rank-3 F32 with equal odd taps <= 15: the single-kernel
        # path, no workspace; REFERENCE: FP64 per-axis passes through two temporaries
        assert (size.value == 0) == fused
    status = lib.iane_hip_convolve_separable_workspace_bytes(ctypes.byref(src), taps, 3, ctypes.byref(src), 1, 2,
                                                             ctypes.byref(size))
    assert status == _lib.ERR_UNSUPPORTED
    assert b"SPLIT_FP16" in lib.iane_hip_last_error()
