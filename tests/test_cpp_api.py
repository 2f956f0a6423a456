"""Runs the C++ drop-in test (tests/cpp/test_api.cpp): the reference's own test
cases against the header-only include/iane API — host path here, device path on the GPU."""
from __future__ import annotations

import subprocess

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
