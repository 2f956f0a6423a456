"""Build the example programs (examples/*/main.cpp) against include/ and libiane_hip.so.

    python examples/build.py        # -> examples/_bin/{convolution,rl_deconvolution}
"""
from __future__ import annotations

import pathlib
import subprocess

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parent
BIN = HERE / "_bin"
PROGRAMS = ("convolution", "rl_deconvolution")


def build(force: bool = False) -> dict:
    BIN.mkdir(exist_ok=True)
    headers = [*sorted((ROOT / "include" / "iane").glob("*.hpp")), ROOT / "include" / "iane_hip.h",
               *sorted((HERE / "common").glob("*.hpp"))]
    out = {}
    for name in PROGRAMS:
        source = HERE / name / "main.cpp"
        binary = BIN / name
        stale = force or not binary.exists() or any(binary.stat().st_mtime < s.stat().st_mtime
                                                    for s in [source, *headers])
        if stale:
            cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-Wextra", "-pthread", f"-I{ROOT / 'include'}", str(source),
                   "-o", str(binary), f"-L{ROOT / 'iane_amd' / 'lib'}", "-liane_hip",
                   "-Wl,-rpath,$ORIGIN/../../iane_amd/lib"]
            subprocess.run(cmd, check=True)
        out[name] = binary
    return out


if __name__ == "__main__":
    for name, path in build(force=True).items():
        print(name, path)
