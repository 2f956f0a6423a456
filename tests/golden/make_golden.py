"""Writes the golden fixtures under tests/golden/ from the reference itself.

Run where the reference headers exist (oracle/_ref/libiane_ref.so built by
`make -C oracle`):  python tests/golden/make_golden.py

Every fixture is data: the inputs and the reference's outputs.
  conv_*.npz : src, kernel, expected (reference iane::convolve with cpu_policy), border
  rl_*.npz   : observed, psf, iterations, epsilon, expected (the RL loop of
               examples/rl_deconvolution/main.cu:51-81 run through the reference's
               convolve / zip_view / copy on float images)
"""
from __future__ import annotations

import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from iane_amd import synth  # noqa: E402
from oracle import oracle  # noqa: E402

OUT = pathlib.Path(__file__).resolve().parent
DT = {"u8": np.uint8, "u16": np.uint16, "i32": np.int32, "f32": np.float32, "f64": np.float64}


def save_conv(name, src, ker, ddt, border):
    expected = oracle.convolve_ref(src, ker, ddt, border)
    np.savez_compressed(OUT / f"conv_{name}.npz", src=src, kernel=ker, expected=expected,
                        border=np.int32(border))


def main():
    if not oracle.have_reference():
        raise SystemExit("oracle/_ref/libiane_ref.so missing: build it with `make -C oracle` next to the reference")
    rng = np.random.default_rng(2024)
    # dtype matrix at ranks 1..4
    combos = [("u8", "f32", "u8"), ("u8", "i32", "i32"), ("u16", "i32", "i32"), ("i32", "i32", "i32"),
              ("f32", "f32", "f32"), ("f64", "f64", "f64"), ("f32", "f64", "u16"), ("u16", "f32", "f32"),
              ("f32", "f32", "i32"), ("i32", "f64", "u8")]
    shapes = {1: ((37,), (5,)), 2: ((19, 23), (5, 3)), 3: ((9, 11, 13), (3, 5, 3)), 4: ((4, 5, 6, 7), (3, 1, 3, 3))}
    for sdt, kdt, ddt in combos:
        for rank, (shape, kshape) in shapes.items():
            for border in (0, 1):
                s = DT[sdt]
                if np.issubdtype(s, np.integer):
                    hi = {np.uint8: 256, np.uint16: 65536, np.int32: 5000}[s]
                    lo = -5000 if s == np.int32 else 0
                    src = rng.integers(lo, hi, shape).astype(s)
                else:
                    src = (rng.random(shape) * 300 - 20).astype(s)
                k = DT[kdt]
                ker = rng.integers(-4, 5, kshape).astype(k) if np.issubdtype(k, np.integer) else \
                    (rng.random(kshape) - 0.25).astype(k)
                save_conv(f"{sdt}_{kdt}_{ddt}_r{rank}_{'clamp' if border == 0 else 'zero'}", src, ker, DT[ddt], border)
    # the configurations' kernels on small crops of their inputs
    img, gauss7, img_u8, box5 = synth.image_2d_uniform(64, seed=1)
    save_conv("img2d_gauss7_clamp", img, gauss7, np.float32, 0)
    save_conv("img2d_u8_box5_clamp", img_u8, box5, np.int32, 0)
    vol = synth.volume_blobs(24, seed=2, blobs=6)
    save_conv("vol3d_dense15_zero", vol["volume"], vol["dense"], np.float32, 1)
    taps = vol["taps"].astype(np.float64)
    sep = np.multiply.outer(np.multiply.outer(taps, taps), taps).astype(np.float32)
    save_conv("vol3d_gauss15_zero", vol["volume"], sep, np.float32, 1)
    save_conv("vol3d_u16_binomial_zero", vol["volume_u16"], vol["binomial"], np.int32, 1)
    # wide kernels (untemplated widths) and even-sized images
    src = rng.random((40, 42)).astype(np.float32)
    save_conv("wide17x19_clamp", src, (rng.random((19, 17)) - 0.5).astype(np.float32), np.float32, 0)
    save_conv("wide33x1_zero", src, (rng.random((1, 33)) - 0.5).astype(np.float32), np.float32, 1)
    # RL loop
    for name, shape, psf_sigma, psf_size, iters in [("2d", (24, 20), (1.2, 1.0), (5, 5), 10),
                                                     ("3d", (10, 14, 12), (1.0, 1.2, 0.8), (3, 5, 5), 6)]:
        truth = np.full(shape, 0.05)
        synth.add_blobs(truth, 4, (1.0, 2.5), (0.5, 1.0), 7, 1)
        psf = synth.gaussian_kernel(psf_sigma, psf_size).astype(np.float32)
        blurred = oracle.convolve(truth * 200, psf.astype(np.float64), np.float64, oracle.ZERO)
        observed = synth.poisson(np.maximum(blurred, 0), 7, 2).reshape(shape).astype(np.float32)
        expected = oracle.rl_ref(observed, psf, iters, 1e-6)
        np.savez_compressed(OUT / f"rl_{name}.npz", observed=observed, psf=psf, iterations=np.int32(iters),
                            epsilon=np.float32(1e-6), expected=expected)
    print("fixtures written to", OUT)


if __name__ == "__main__":
    main()
