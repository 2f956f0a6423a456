"""Compile the C++ drop-in test (tests/cpp/test_api.cpp) against include/ and libiane_hip.so."""
from __future__ import annotations

import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent.parent
BIN = HERE / "_bin" / "test_api"


def build(force: bool = False) -> pathlib.Path:
    sources = [HERE / "test_api.cpp", *sorted((ROOT / "include" / "iane").glob("*.hpp")), ROOT / "include" / "iane_hip.h"]
    if not force and BIN.exists() and all(BIN.stat().st_mtime >= s.stat().st_mtime for s in sources):
        return BIN
    BIN.parent.mkdir(exist_ok=True)
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-pthread", f"-I{ROOT / 'include'}", str(HERE / "test_api.cpp"),
           "-o", str(BIN), f"-L{ROOT / 'iane_amd' / 'lib'}", "-liane_hip", "-Wl,-rpath,$ORIGIN/../../../iane_amd/lib"]
    subprocess.run(cmd, check=True)
    return BIN


if __name__ == "__main__":
    print(build(force=True))
