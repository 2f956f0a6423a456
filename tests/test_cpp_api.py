"""Runs the C++ drop-in test (tests/cpp/test_api.cpp): the reference's own test
cases against the header-only include/iane API — host path here, device path on the GPU."""
from __future__ import annotations

import pathlib
import subprocess

import numpy as np
import pytest

from tests.cpp import build as cpp_build


def _run(mode: str) -> subprocess.CompletedProcess:
    binary = cpp_build.build()
    return subprocess.run([str(binary), mode], capture_output=True, text=True, timeout=600)


def test_cpp_api_host_path():
    proc = _run("host")
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "0 failures" in proc.stdout


@pytest.mark.gpu
def test_cpp_api_device_path():
    proc = _run("device")
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "0 failures" in proc.stdout


GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def _rl_golden(tmp_path, name: str, mode: str) -> subprocess.CompletedProcess:
    """The reference RL loop's fixture as raw files for the C++ test, which runs
    the loop written with constant_view / transform / zip_view / copy."""
    data = np.load(GOLDEN / f"{name}.npz", allow_pickle=False)
    out = tmp_path / name
    out.mkdir()
    for key in ("observed", "psf", "expected"):
        np.ascontiguousarray(data[key], dtype=np.float32).tofile(out / f"{key}.f32")
    shape, kshape = data["observed"].shape[::-1], data["psf"].shape[::-1]
    (out / "meta.txt").write_text(f"{len(shape)} {' '.join(map(str, shape))} {' '.join(map(str, kshape))} "
                                  f"{int(data['iterations'])} {float(data['epsilon']):.9g}\n")
    args = [str(cpp_build.build()), "rl-golden", str(out)] + (["device"] if mode == "device" else [])
    return subprocess.run(args, capture_output=True, text=True, timeout=600)


@pytest.mark.parametrize("name", ["rl_2d", "rl_3d"])
def test_cpp_rl_written_like_the_reference_host(tmp_path, name):
    proc = _rl_golden(tmp_path, name, "host")
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "0 failures" in proc.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["rl_2d", "rl_3d"])
def test_cpp_rl_written_like_the_reference_device(tmp_path, name):
    proc = _rl_golden(tmp_path, name, "device")
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "0 failures" in proc.stdout
