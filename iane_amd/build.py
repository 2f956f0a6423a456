"""Build libiane_hip.so in-tree: every HIP source under csrc/ compiled for gfx950.

`python -m iane_amd.build` or `__graft_entry__.build()`. Object files go to
iane_amd/_build/, the shared library to iane_amd/lib/libiane_hip.so.
"""
from __future__ import annotations

import concurrent.futures as futures
import os
import pathlib
import subprocess
import sys

ROOT = pathlib.Path(__file__).resolve().parent
CSRC = ROOT / "csrc"
BUILD = ROOT / "_build"
LIB = ROOT / "lib" / "libiane_hip.so"
INCLUDE = ROOT.parent / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are disabled for all device code (some of their
# forms have produced wrong results on the target boxes). Contraction into
# fma is decided per source expression (-ffp-contract=on), so a device
# function inlined into different kernels (single-volume and batched passes,
# tuned launch shapes) performs exactly the same arithmetic.
FLAGS = [
    "-std=c++20", "-O3", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=on",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", f"-I{INCLUDE}",
]


# Per-source additions (measured on the separable 15^3 headline):
#   conv_separable: the iterative ILP scheduling strategy (FP32 engine 57.6 ->
#     50.4 us: it interleaves the x, y and z work of neighbouring planes);
#   conv_separable_split: matrix-instruction results in VGPRs (every filter's
#     accumulators feed the next filter's operand split directly);
#   conv_mfma: a higher full-unroll limit, so the split kernel's rows of one
#     pass (up to 30 rows of 45 matrix instructions) unroll and its operand
#     and row buffers stay in registers.
EXTRA_FLAGS = {
    "conv_separable": ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"],
    "conv_separable_split": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
    "conv_mfma": ["-mllvm", "-pragma-unroll-threshold=200000"],
}


def _compile(src: pathlib.Path) -> pathlib.Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *FLAGS, *EXTRA_FLAGS.get(src.stem, []), "-c", str(src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{proc.stdout}\n{proc.stderr}")
    return obj


def build(verbose: bool = True) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    LIB.parent.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    jobs = min(len(sources), max(1, (os.cpu_count() or 4)), 8)
    with futures.ThreadPoolExecutor(jobs) as pool:
        objects = list(pool.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objects)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed:\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(sources)} sources")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
