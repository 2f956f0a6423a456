"""The example programs (examples/convolution, examples/rl_deconvolution): the
reference's CLIs (examples/common/cli.{hpp,cpp}, kernel.hpp, convolution/main.cu,
rl_deconvolution/main.cu) on this backend.

CPU tests pin the PGM/PPM reader and writer against numpy-written files, the
convolution output against the oracle (byte-exact, reference arithmetic),
the RL output against the reference's own loop compiled from its headers
(oracle/_ref), the JSON configuration rules and the error messages. GPU tests
check that the hip backend writes the same bytes as the cpu backend.
"""
from __future__ import annotations

import json
import pathlib
import subprocess

import numpy as np
import pytest

from oracle import oracle
from tests import cases

ROOT = pathlib.Path(__file__).resolve().parent.parent


def write_pnm(path, image: np.ndarray, maxval: int = 255, comment: bool = False) -> None:
    """Binary PGM (2-D array) or PPM (H x W x 3), written here independently of the C++ code."""
    magic = b"P5" if image.ndim == 2 else b"P6"
    h, w = image.shape[:2]
    header = magic + (b"\n# made by the test\n" if comment else b"\n") + f"{w} {h}\n{maxval}\n".encode()
    pathlib.Path(path).write_bytes(header + np.ascontiguousarray(image, dtype=np.uint8).tobytes())


def read_pnm(path) -> np.ndarray:
    data = pathlib.Path(path).read_bytes()
    assert data[:2] in (b"P5", b"P6")
    fields, pos = [], 2
    while len(fields) < 3:
        while data[pos:pos + 1].isspace():
            pos += 1
        start = pos
        while data[pos:pos + 1].isdigit():
            pos += 1
        fields.append(int(data[start:pos]))
    pos += 1
    w, h, maxval = fields
    assert maxval == 255
    channels = 1 if data[:2] == b"P5" else 3
    pixels = np.frombuffer(data[pos:pos + w * h * channels], dtype=np.uint8)
    return pixels.reshape((h, w) if channels == 1 else (h, w, 3))


@pytest.fixture(scope="module")
def programs():
    import sys
    sys.path.insert(0, str(ROOT / "examples"))
    import build  # examples/build.py
    return build.build()


def run(binary, *args, cwd):
    return subprocess.run([str(binary), *map(str, args)], capture_output=True, text=True, cwd=cwd, timeout=300)


def to_gray8(v: np.ndarray) -> np.ndarray:
    v = v.astype(np.float32)
    whole = np.floor(v)
    rounded = whole + ((v - whole) >= np.float32(0.5))
    return np.where(~(v > 0), 0, np.where(v >= 255, 255, rounded)).astype(np.uint8)


IDENTITY = "1 1\n1\n"


# ---------------------------------------------------------------- codecs ----
@pytest.mark.parametrize("channels,maxval,comment", [(1, 255, False), (3, 255, True), (1, 100, True), (3, 15, False)])
def test_pnm_round_trip(programs, tmp_path, channels, maxval, comment):
    """Binary PGM/PPM in, the identity kernel, the same pixels out (maxval below 255
    scaled to 0..255 with rounding; header comments skipped)."""
    rng = np.random.default_rng(4 + channels + maxval)
    data = rng.integers(0, maxval + 1, (9, 14, 3) if channels == 3 else (9, 14), dtype=np.uint8)
    ext = ".pgm" if channels == 1 else ".ppm"
    write_pnm(tmp_path / f"in{ext}", data, maxval, comment)
    (tmp_path / "k.txt").write_text(IDENTITY)
    proc = run(programs["convolution"], f"in{ext}", f"out{ext}", "--kernel", "k.txt", "--runs", 1, "--warmup", 0,
               cwd=tmp_path)
    assert proc.returncode == 0, proc.stderr
    expected = ((data.astype(np.int64) * 255 + maxval // 2) // maxval).astype(np.uint8)
    np.testing.assert_array_equal(read_pnm(tmp_path / f"out{ext}"), expected)


def test_other_formats_are_refused_clearly(programs, tmp_path):
    (tmp_path / "k.txt").write_text(IDENTITY)
    (tmp_path / "in.png").write_bytes(b"\x89PNG\r\n\x1a\n" + bytes(32))
    proc = run(programs["convolution"], "in.png", "out.pgm", "--kernel", "k.txt", cwd=tmp_path)
    assert proc.returncode == 1 and "unsupported image format (binary PGM or PPM expected)" in proc.stderr
    write_pnm(tmp_path / "in.pgm", np.zeros((4, 4), np.uint8))
    proc = run(programs["convolution"], "in.pgm", "out.png", "--kernel", "k.txt", cwd=tmp_path)
    assert proc.returncode == 1 and "output must be a binary .pgm (grey) or .ppm (colour) file" in proc.stderr
    proc = run(programs["convolution"], "in.pgm", "out.ppm", "--kernel", "k.txt", cwd=tmp_path)
    assert proc.returncode == 1 and ".pgm holds grey images and .ppm colour images" in proc.stderr


# ----------------------------------------------------------- convolution ----
@pytest.mark.parametrize("border", ["clamp", "zero"])
def test_grey_convolution_with_roi_matches_oracle(programs, tmp_path, border):
    rng = np.random.default_rng(5)
    image = rng.integers(0, 256, (40, 57), dtype=np.uint8)
    write_pnm(tmp_path / "in.pgm", image)
    kernel = (rng.random((3, 5)) - 0.2).astype(np.float32)
    (tmp_path / "k.txt").write_text("5 3\n" + "\n".join(" ".join(repr(float(v)) for v in row) for row in kernel))
    proc = run(programs["convolution"], "in.pgm", "out.pgm", "--kernel", "k.txt", "--roi", 16, 8, 30, 21,
               "--border", border, "--runs", 2, "--warmup", 1, cwd=tmp_path)
    assert proc.returncode == 0, proc.stderr
    assert "backend=cpu size=30x21 warmup=1 runs=2" in proc.stdout and "compute_ms min=" in proc.stdout
    crop = np.ascontiguousarray(image[8:29, 16:46])
    expected = oracle.convolve(crop, kernel, np.uint8, cases.CLAMP if border == "clamp" else cases.ZERO)
    np.testing.assert_array_equal(read_pnm(tmp_path / "out.pgm"), expected)


def test_colour_channels_convolve_independently(programs, tmp_path):
    rng = np.random.default_rng(6)
    image = rng.integers(0, 256, (25, 31, 3), dtype=np.uint8)
    write_pnm(tmp_path / "in.ppm", image)
    (tmp_path / "k.txt").write_text("3 3\n1 2 1\n2 4 2\n1 2 1\n")
    proc = run(programs["convolution"], "in.ppm", "out.ppm", "--kernel", "k.txt", "--roi", 2, 3, 20, 15,
               "--border", "zero", "--runs", 1, cwd=tmp_path)
    assert proc.returncode == 0, proc.stderr
    kernel = np.array([[1, 2, 1], [2, 4, 2], [1, 2, 1]], np.float32)
    got = read_pnm(tmp_path / "out.ppm")
    for c in range(3):
        crop = np.ascontiguousarray(image[3:18, 2:22, c])
        np.testing.assert_array_equal(got[..., c], oracle.convolve(crop, kernel, np.uint8, cases.ZERO))


def test_orientation_reads_the_pixel_to_the_right(programs, tmp_path):
    """README example: the 3x3 kernel with 1 at (x=0, y=1) makes an output read its right neighbour."""
    image = np.arange(48, dtype=np.uint8).reshape(6, 8) * 5
    write_pnm(tmp_path / "in.pgm", image)
    (tmp_path / "k.txt").write_text("3 3\n0 0 0\n1 0 0\n0 0 0\n")
    assert run(programs["convolution"], "in.pgm", "out.pgm", "--kernel", "k.txt", "--runs", 1,
               cwd=tmp_path).returncode == 0
    out = read_pnm(tmp_path / "out.pgm")
    np.testing.assert_array_equal(out[:, :-1], image[:, 1:])


# --------------------------------------------------------- configuration ----
def test_json_configuration_and_command_line_override(programs, tmp_path):
    image = np.random.default_rng(7).integers(0, 256, (30, 40), dtype=np.uint8)
    write_pnm(tmp_path / "in.pgm", image)
    (tmp_path / "k.txt").write_text(IDENTITY)
    (tmp_path / "settings.json").write_text(json.dumps(
        {"input": "in.pgm", "output": "out.pgm", "kernel": "k.txt", "roi": [4, 5, 10, 6], "runs": 1, "warmup": 0}))
    proc = run(programs["convolution"], "--config", "settings.json", cwd=tmp_path)
    assert proc.returncode == 0, proc.stderr
    np.testing.assert_array_equal(read_pnm(tmp_path / "out.pgm"), image[5:11, 4:14])
    proc = run(programs["convolution"], "--config", "settings.json", "--roi", 1, 2, 3, 4, cwd=tmp_path)
    assert proc.returncode == 0, proc.stderr
    np.testing.assert_array_equal(read_pnm(tmp_path / "out.pgm"), image[2:6, 1:4])
    (tmp_path / "bad.json").write_text(json.dumps({"input": "in.pgm", "output": "o.pgm", "kernel": "k.txt",
                                                   "colour": 1}))
    proc = run(programs["convolution"], "--config", "bad.json", cwd=tmp_path)
    assert proc.returncode == 1 and "Unknown configuration option 'colour'" in proc.stderr
    (tmp_path / "bad2.json").write_text(json.dumps({"input": "in.pgm", "output": "o.pgm", "kernel": "k.txt",
                                                    "border": ["clamp", "zero"]}))
    proc = run(programs["convolution"], "--config", "bad2.json", cwd=tmp_path)
    assert proc.returncode == 1 and "Option 'border': Wrong number of values" in proc.stderr


@pytest.mark.parametrize("kernel_text,message", [
    ("2 3\n1 1 1 1 1 1\n", "Expected positive odd kernel dimensions"),
    ("3 1\n1 nan 1\n", "Expected finite kernel coefficients"),
    ("3 1\n1 2\n", "Expected finite kernel coefficients"),
    ("1 1\n1 2\n", "Unexpected data after kernel coefficients"),
])
def test_kernel_file_errors(programs, tmp_path, kernel_text, message):
    write_pnm(tmp_path / "in.pgm", np.zeros((4, 4), np.uint8))
    (tmp_path / "k.txt").write_text(kernel_text)
    proc = run(programs["convolution"], "in.pgm", "out.pgm", "--kernel", "k.txt", cwd=tmp_path)
    assert proc.returncode == 1 and proc.stderr.startswith("convolution: ") and message in proc.stderr


@pytest.mark.parametrize("args,message", [
    (["--roi", 0, 0, 5, 4], "ROI must be a nonempty rectangle inside the input"),
    (["--roi", 1, 2, 3], "ROI must be a nonempty rectangle inside the input"),
    (["--border", "wrap"], "Border must be clamp or zero"),
    (["--backend", "opencl"], "Backend must be cpu or hip"),
    (["--runs", 0], "runs must be positive"),
    (["--runs", "x"], "for option '--runs' is invalid"),
    (["--colour", 1], "unrecognised option '--colour'"),
])
def test_option_errors(programs, tmp_path, args, message):
    write_pnm(tmp_path / "in.pgm", np.zeros((4, 4), np.uint8))
    (tmp_path / "k.txt").write_text(IDENTITY)
    proc = run(programs["convolution"], "in.pgm", "out.pgm", "--kernel", "k.txt", *args, cwd=tmp_path)
    assert proc.returncode == 1 and message in proc.stderr, proc.stderr


# -------------------------------------------------------------------- RL ----
def _rl_case(tmp_path):
    rng = np.random.default_rng(9)
    y, x = np.mgrid[0:28, 0:36]
    truth = 20 + 180 * np.exp(-((x - 12) ** 2 + (y - 10) ** 2) / 18.0) + 120 * np.exp(-((x - 26) ** 2 + (y - 19) ** 2) / 8.0)
    image = np.clip(rng.poisson(truth), 0, 255).astype(np.uint8)
    write_pnm(tmp_path / "blurred.pgm", image)
    weights = np.outer([1, 2, 1], [1, 4, 6, 4, 1]).astype(np.float64)
    (tmp_path / "psf.txt").write_text("5 3\n" + "\n".join(" ".join(str(int(v)) for v in row) for row in weights))
    psf = (weights / weights.sum()).astype(np.float32)   # read_psf: double sum, stored as float
    return image, psf


@pytest.mark.skipif(not oracle.have_reference(), reason="the reference's own RL loop (oracle/_ref) is not built")
def test_rl_cpu_backend_equals_reference_loop(programs, tmp_path):
    image, psf = _rl_case(tmp_path)
    proc = run(programs["rl_deconvolution"], "blurred.pgm", "restored.pgm", "--psf", "psf.txt", "--iterations", 12,
               "--runs", 1, "--warmup", 0, cwd=tmp_path)
    assert proc.returncode == 0, proc.stderr
    assert "prepare_ms=" in proc.stdout and "compute_ms min=" in proc.stdout
    expected = to_gray8(oracle.rl_ref(image.astype(np.float32), psf, 12, 1e-6))
    np.testing.assert_array_equal(read_pnm(tmp_path / "restored.pgm"), expected)


def test_rl_option_errors(programs, tmp_path):
    _rl_case(tmp_path)
    for args, message in [(["--iterations", 0], "Iterations and finite epsilon must be positive"),
                          (["--epsilon", -1], "Iterations and finite epsilon must be positive"),
                          (["--method", "wiener"], "Method must be direct or fft")]:
        proc = run(programs["rl_deconvolution"], "blurred.pgm", "r.pgm", "--psf", "psf.txt", *args, cwd=tmp_path)
        assert proc.returncode == 1 and message in proc.stderr, proc.stderr
    (tmp_path / "neg.txt").write_text("3 1\n1 -1 1\n")
    proc = run(programs["rl_deconvolution"], "blurred.pgm", "r.pgm", "--psf", "neg.txt", cwd=tmp_path)
    assert proc.returncode == 1 and "PSF coefficients must be nonnegative" in proc.stderr
    write_pnm(tmp_path / "colour.ppm", np.zeros((4, 4, 3), np.uint8))
    proc = run(programs["rl_deconvolution"], "colour.ppm", "r.pgm", "--psf", "psf.txt", cwd=tmp_path)
    assert proc.returncode == 1 and "expected a grayscale image" in proc.stderr


# ------------------------------------------------------------------- GPU ----
@pytest.mark.gpu
@pytest.mark.parametrize("border", ["clamp", "zero"])
def test_hip_backend_writes_the_cpu_bytes(programs, tmp_path, border):
    rng = np.random.default_rng(10)
    write_pnm(tmp_path / "g.pgm", rng.integers(0, 256, (70, 90), dtype=np.uint8))
    write_pnm(tmp_path / "c.ppm", rng.integers(0, 256, (41, 53, 3), dtype=np.uint8))
    (tmp_path / "k.txt").write_text("7 5\n" + "\n".join(" ".join(repr(float(v)) for v in row)
                                                     for row in (rng.random((5, 7)) - 0.3)))
    for name, ext, roi in (("g", ".pgm", [10, 7, 60, 50]), ("c", ".ppm", [3, 4, 40, 30])):
        outs = []
        for backend in ("cpu", "hip"):
            out = f"{name}_{backend}{ext}"
            proc = run(programs["convolution"], f"{name}{ext}", out, "--kernel", "k.txt", "--roi", *roi,
                       "--border", border, "--backend", backend, "--runs", 2, "--warmup", 1, cwd=tmp_path)
            assert proc.returncode == 0, proc.stderr
            outs.append(read_pnm(tmp_path / out))
        assert "backend=hip device=0" in proc.stdout and "upload_ms=" in proc.stdout and "download_ms=" in proc.stdout
        np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.gpu
def test_rl_hip_direct_equals_cpu_and_fft_is_close(programs, tmp_path):
    _rl_case(tmp_path)
    got = {}
    for backend, method in (("cpu", "direct"), ("hip", "direct"), ("hip", "fft")):
        out = f"r_{backend}_{method}.pgm"
        proc = run(programs["rl_deconvolution"], "blurred.pgm", out, "--psf", "psf.txt", "--iterations", 15,
                   "--backend", backend, "--method", method, "--runs", 1, "--warmup", 0, cwd=tmp_path)
        assert proc.returncode == 0, proc.stderr
        got[(backend, method)] = read_pnm(tmp_path / out).astype(int)
    np.testing.assert_array_equal(got[("cpu", "direct")], got[("hip", "direct")])
    # the spectral method sums in another order: bytes may differ by one rounding step
    assert np.abs(got[("hip", "fft")] - got[("cpu", "direct")]).max() <= 1
