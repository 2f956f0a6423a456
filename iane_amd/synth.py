"""Seeded synthetic workloads for the benchmarks and tests (host side, numpy).

Counter-based Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy
as 1, 2, 3", SC'11): key = (seed, stream), counter = (index_lo, index_hi, 0, 0),
four 32-bit words per counter. uniform = (u32 >> 8) * 2^-24; normal by
Box–Muller; Poisson by inversion for lambda < 10 and PTRS (Hörmann 1993)
otherwise. Every array here is a numpy array in C order (IANE axis 0 = last axis).

Workloads (named by what they are):
  * image_2d_uniform       1024^2 U[0,1), 7x7 Gaussian sigma 1.5 (clamp) / u8 + 5x5 box
  * volume_blobs           256^3 U[0,1) + 200 Gaussian blobs, 15^3 Gaussian sigma 2.5,
                           dense 15^3 normal kernel normalised to sum 1, u16 + 3^3 binomial,
                           dense 3^3 normal kernel normalised to sum 1
  * rl_volume              256x256x128 truth (0.01 + 300 blobs) blurred by a 31x31x63
                           anisotropic Gaussian PSF, peak 1000 photons, Poisson noise
"""
from __future__ import annotations

import numpy as np

_M0, _M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
_W0, _W1 = np.uint32(0x9E3779B9), np.uint32(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32(seed: int, stream: int, n_blocks: int, offset: int = 0, counters=None) -> np.ndarray:
    """``n_blocks`` Philox4x32-10 outputs (shape (n_blocks, 4), uint32) for the
    counters offset .. offset + n_blocks - 1, or for the given ``counters``."""
    idx = np.arange(offset, offset + n_blocks, dtype=np.uint64) if counters is None else \
        np.asarray(counters, dtype=np.uint64)
    n_blocks = idx.size
    c0 = (idx & _MASK).astype(np.uint32)
    c1 = (idx >> np.uint64(32)).astype(np.uint32)
    c2 = np.zeros(n_blocks, dtype=np.uint32)
    c3 = np.zeros(n_blocks, dtype=np.uint32)
    k0 = np.uint32(seed & 0xFFFFFFFF)
    k1 = np.uint32(stream & 0xFFFFFFFF)
    with np.errstate(over="ignore"):
        for _ in range(10):
            p0 = _M0 * c0.astype(np.uint64)
            p1 = _M1 * c2.astype(np.uint64)
            hi0, lo0 = (p0 >> np.uint64(32)).astype(np.uint32), (p0 & _MASK).astype(np.uint32)
            hi1, lo1 = (p1 >> np.uint64(32)).astype(np.uint32), (p1 & _MASK).astype(np.uint32)
            c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
            k0 = np.uint32(k0 + _W0)
            k1 = np.uint32(k1 + _W1)
    return np.stack([c0, c1, c2, c3], axis=1)


def uniform(seed: int, stream: int, n: int, offset: int = 0) -> np.ndarray:
    """n float64 values in [0, 1) with 24-bit resolution (exact in float32)."""
    blocks = (n + 3) // 4
    words = philox4x32(seed, stream, blocks, offset).reshape(-1)[:n]
    return (words >> np.uint32(8)).astype(np.float64) * (1.0 / 16777216.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Box–Muller over consecutive uniform pairs."""
    m = (n + 1) // 2
    u = uniform(seed, stream, 2 * m)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
    out = np.empty(2 * m)
    out[0::2] = z[:m]
    out[1::2] = z[m:]
    return out[:n]


def poisson(lam: np.ndarray, seed: int, stream: int) -> np.ndarray:
    """Poisson variates: inversion for lambda < 10, PTRS (Hörmann 1993) otherwise.

    Sample i uses the uniforms of Philox counter block (attempt * n + i)."""
    lam = np.asarray(lam, dtype=np.float64).reshape(-1)
    n = lam.size
    out = np.zeros(n, dtype=np.float64)
    small = lam < 10.0
    # inversion: one uniform, sequential CDF search (vectorised over samples)
    if small.any():
        ls = lam[small]
        u = uniform(seed, stream, small.sum())
        k = np.zeros(ls.size)
        p = np.exp(-ls)
        cdf = p.copy()
        active = u > cdf
        it = 0
        while active.any() and it < 1000:
            k[active] += 1
            p[active] *= ls[active] / k[active]
            cdf[active] += p[active]
            active = u > cdf
            it += 1
        out[small] = k
    big = ~small
    if big.any():
        lb = lam[big]
        m = lb.size
        slam = np.sqrt(lb)
        loglam = np.log(lb)
        b = 0.931 + 2.53 * slam
        a = -0.059 + 0.02483 * b
        inv_alpha = 1.1239 + 1.1328 / (b - 3.4)
        v_r = 0.9277 - 3.6224 / (b - 2)
        result = np.full(m, -1.0)
        attempt = 0
        from scipy.special import gammaln as lg
        idx = np.arange(m)
        while idx.size:
            # sample i, attempt t uses counter block t*m + i (only pending samples are drawn)
            blocks = philox4x32(seed, stream + 1, 0, counters=attempt * m + idx)
            u = (blocks[:, 0] >> np.uint32(8)).astype(np.float64) / 16777216.0 - 0.5
            v = (blocks[:, 1] >> np.uint32(8)).astype(np.float64) / 16777216.0
            us = 0.5 - np.abs(u)
            A, B, IA, VR, L, LL = a[idx], b[idx], inv_alpha[idx], v_r[idx], lb[idx], loglam[idx]
            with np.errstate(divide="ignore", invalid="ignore"):
                k = np.floor((2 * A / us + B) * u + L + 0.43)
                quick = (us >= 0.07) & (v <= VR)
                ok_range = ~((k < 0) | ((us < 0.013) & (v > us)))
                lhs = np.log(v * IA / (A / (us * us) + B))
                rhs = -L + k * LL - lg(np.maximum(k, 0) + 1)
                accept = quick | (ok_range & (lhs <= rhs))
            result[idx[accept]] = k[accept]
            idx = idx[~accept]
            attempt += 1
            if attempt > 200:
                raise RuntimeError("PTRS did not converge")
        out[big] = result
    return out


def gaussian_1d(sigma: float, size: int) -> np.ndarray:
    x = np.arange(size) - size // 2
    g = np.exp(-(x * x) / (2.0 * sigma * sigma))
    return g / g.sum()


def gaussian_kernel(sigmas, sizes) -> np.ndarray:
    """Normalised separable Gaussian; ``sigmas``/``sizes`` in array-axis order."""
    k = np.ones(())
    for s, n in zip(sigmas, sizes):
        k = np.multiply.outer(k, gaussian_1d(s, n))
    return k / k.sum()


def add_blobs(volume: np.ndarray, count: int, sigma_range, amp_range, seed: int, stream: int) -> None:
    """Add ``count`` isotropic Gaussian blobs (centre uniform in the volume) in place."""
    u = uniform(seed, stream, count * (volume.ndim + 2))
    u = u.reshape(count, volume.ndim + 2)
    for row in u:
        centre = [row[i] * volume.shape[i] for i in range(volume.ndim)]
        sigma = sigma_range[0] + row[volume.ndim] * (sigma_range[1] - sigma_range[0])
        amp = amp_range[0] + row[volume.ndim + 1] * (amp_range[1] - amp_range[0])
        r = int(np.ceil(4 * sigma))
        sl = []
        grids = []
        for i, c in enumerate(centre):
            lo = max(0, int(np.floor(c)) - r)
            hi = min(volume.shape[i], int(np.floor(c)) + r + 1)
            sl.append(slice(lo, hi))
            grids.append((np.arange(lo, hi) + 0.5 - c) ** 2)
        d2 = grids[0]
        for g in grids[1:]:
            d2 = np.add.outer(d2, g)
        volume[tuple(sl)] += amp * np.exp(-d2 / (2 * sigma * sigma))


# ------------------------------------------------------------- workloads ---
def image_2d_uniform(size: int = 1024, seed: int = 1):
    img = uniform(seed, 0, size * size).reshape(size, size).astype(np.float32)
    kernel = gaussian_kernel((1.5, 1.5), (7, 7)).astype(np.float32)
    img_u8 = np.minimum(255, np.floor(img.astype(np.float64) * 256)).astype(np.uint8)
    box = np.ones((5, 5), dtype=np.int32)
    return img, kernel, img_u8, box


def volume_blobs(n: int = 256, seed: int = 2, blobs: int = 200):
    vol = uniform(seed, 0, n ** 3).reshape(n, n, n)
    add_blobs(vol, blobs, (2.0, 8.0), (0.5, 2.0), seed, 1)
    vol = vol.astype(np.float32)
    taps = gaussian_1d(2.5, 15)
    dense = normal(seed, 2, 15 ** 3).reshape(15, 15, 15)
    dense_sum = float(dense.sum())
    dense = (dense / dense_sum).astype(np.float32)
    vol_u16 = np.minimum(65535, np.floor(vol.astype(np.float64) * 8192)).astype(np.uint16)
    b = np.array([1, 2, 1], dtype=np.int32)
    binomial = np.multiply.outer(np.multiply.outer(b, b), b).astype(np.int32)
    dense3 = normal(seed, 3, 27).reshape(3, 3, 3)
    dense3 = (dense3 / float(dense3.sum())).astype(np.float32)
    return {"volume": vol, "taps": taps.astype(np.float32), "dense": dense, "dense_sum": dense_sum,
            "volume_u16": vol_u16, "binomial": binomial, "dense3": dense3}


def rl_batch_data(count: int = 64, seed: int = 4, edge_range=(32, 192), sigma_range=(1.0, 4.0),
                  peak_range=(200.0, 2000.0), blobs_per_megavoxel: float = 150.0):
    """``count`` RL problems of uneven shapes: each axis length U{32..192}, its own
    anisotropic Gaussian PSF (sigma U[1,4] per axis, support 2*ceil(3 sigma)+1),
    Poisson noise at a peak U[200, 2000] photons. Returns a list of
    (observed float32, psf float32) in array order (z, y, x)."""
    u = uniform(seed, 0, count * 7).reshape(count, 7)
    out = []
    for v in range(count):
        shape = tuple(int(edge_range[0] + np.floor(u[v, a] * (edge_range[1] - edge_range[0] + 1))) for a in range(3))
        sigma = tuple(sigma_range[0] + u[v, 3 + a] * (sigma_range[1] - sigma_range[0]) for a in range(3))
        size = tuple(2 * int(np.ceil(3 * s)) + 1 for s in sigma)
        peak = peak_range[0] + u[v, 6] * (peak_range[1] - peak_range[0])
        truth = np.full(shape, 0.01)
        add_blobs(truth, max(4, int(blobs_per_megavoxel * np.prod(shape) / 1e6)), (1.0, 4.0), (0.1, 1.0),
                  seed * 1000 + v, 1)
        psf = gaussian_kernel(sigma, size)
        blurred = _fftconvolve_same(truth, psf)
        blurred *= peak / blurred.max()
        observed = poisson(np.maximum(blurred, 0.0), seed * 1000 + v, 2).reshape(shape)
        out.append((observed.astype(np.float32), psf.astype(np.float32)))
    return out


def _fftconvolve_same(x: np.ndarray, k: np.ndarray) -> np.ndarray:
    shape = [n + m - 1 for n, m in zip(x.shape, k.shape)]
    axes = list(range(x.ndim))
    full = np.fft.irfftn(np.fft.rfftn(x, shape, axes) * np.fft.rfftn(k, shape, axes), shape, axes)
    return full[tuple(slice(m // 2, m // 2 + n) for n, m in zip(x.shape, k.shape))]


def rl_volume_data(shape=(128, 256, 256), psf_sigma=(3.0, 1.5, 1.5), psf_size=(63, 31, 31), peak=1000.0,
                   seed: int = 3, blobs: int = 300):
    """Truth = 0.01 + blobs (sigma U[1,4], amplitude U[0.1,1]); b = H truth (zero
    border) scaled to ``peak``; observed = Poisson(b). Returns float32 arrays
    (truth_scaled, psf, observed)."""
    truth = np.full(shape, 0.01)
    add_blobs(truth, blobs, (1.0, 4.0), (0.1, 1.0), seed, 1)
    psf = gaussian_kernel(psf_sigma, psf_size)
    blurred = _fftconvolve_same(truth, psf)
    scale = peak / blurred.max()
    blurred *= scale
    observed = poisson(np.maximum(blurred, 0.0), seed, 2).reshape(shape)
    return (truth * scale).astype(np.float32), psf.astype(np.float32), observed.astype(np.float32)
