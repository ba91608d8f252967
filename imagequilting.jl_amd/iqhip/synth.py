"""Seeded synthetic training images for the benchmark workloads.

Gaussian random fields use the exponential covariance C(h) = exp(-3 |h / a|)
(practical-range convention: the correlation drops to 5% at lag a, per axis
ranges a = (a_x, a_y, a_z)), sampled exactly by circulant embedding on a
doubled periodic grid, then standardised to mean 0 and variance 1.
Seeds default to 20261202 + workload number.
"""
from __future__ import annotations

import numpy as np

SEED0 = 20261202


def grf(shape, ranges, rng: np.random.Generator, dtype=np.float32) -> np.ndarray:
    shape = tuple(int(s) for s in shape)
    ranges = tuple(float(r) for r in ranges)
    big = tuple(2 * s for s in shape)
    axes = [np.minimum(np.arange(n), n - np.arange(n)) / r for n, r in zip(big, ranges)]
    h2 = np.zeros(big)
    for d, a in enumerate(axes):
        sh = [1] * len(big)
        sh[d] = -1
        h2 = h2 + (a.reshape(sh)) ** 2
    cov = np.exp(-3.0 * np.sqrt(h2))
    lam = np.real(np.fft.fftn(cov))
    lam[lam < 0] = 0.0
    z = rng.standard_normal(big) + 1j * rng.standard_normal(big)
    f = np.fft.ifftn(np.sqrt(lam) * z) * np.sqrt(np.prod(big))
    x = np.real(f)[tuple(slice(0, s) for s in shape)]
    x = (x - x.mean()) / x.std()
    return x.astype(dtype)


def box_blur(x: np.ndarray, width: int) -> np.ndarray:
    from scipy.ndimage import uniform_filter
    return uniform_filter(x.astype(np.float64), size=width, mode="reflect").astype(x.dtype)


def workload(k: int, seed: int | None = None) -> dict:
    """Inputs of benchmark workload k (1-5) as keyword arguments of iqsim."""
    rng = np.random.default_rng(SEED0 + k if seed is None else seed)
    if k == 1:  # 2-D binary Float64 200x200, 30% ones
        x = grf((200, 200), (20, 20), rng, np.float64)
        T = (x > np.quantile(x, 0.70)).astype(np.float64)
        return dict(trainimg=T, tilesize=(30, 30), simsize=(400, 400), overlap=(1 / 6, 1 / 6), tol=0.1, nreal=1)
    if k == 2:  # 3-D continuous Float32 100x100x50
        T = grf((100, 100, 50), (20, 20, 5), rng, np.float32)
        return dict(trainimg=T, tilesize=(20, 20, 5), simsize=(200, 200, 50), overlap=(1 / 5, 1 / 5, 1 / 5),
                    tol=0.1, nreal=1)
    if k == 3:  # 3-facies channels 150x150x30, 64 realizations of varying size
        T = channels((150, 150, 30), rng)
        sizes = [(int(rng.integers(100, 301)), int(rng.integers(100, 301)), 30) for _ in range(64)]
        return dict(trainimg=T, tilesize=(24, 24, 6), simsizes=sizes, overlap=(0.25, 0.25, 0.25), tol=0.1)
    if k == 4:  # 2-D soft + hard
        T = grf((512, 512), (32, 32), rng, np.float32)
        auxti = (box_blur(T, 9) + rng.normal(0, 0.1, T.shape)).astype(np.float32)
        F2 = grf((1024, 1024), (32, 32), rng, np.float32)
        aux = (box_blur(F2, 9) + rng.normal(0, 0.1, F2.shape)).astype(np.float32)
        flat = rng.choice(1024 * 1024, 500, replace=False)
        hard = {(int(i % 1024), int(i // 1024)): float(F2[i % 1024, i // 1024]) for i in flat}
        return dict(trainimg=T, tilesize=(48, 48), simsize=(1024, 1024), overlap=(1 / 6, 1 / 6), tol=0.05,
                    nreal=32, soft=[(aux, auxti)], hard=hard)
    if k == 5:
        T = grf((256, 256, 64), (40, 40, 8), rng, np.float32)
        return dict(trainimg=T, tilesize=(32, 32, 8), simsize=(384, 384, 64), overlap=(0.25, 0.25, 0.25), tol=0.1,
                    nreal=16)
    raise ValueError(k)


def channels(shape, rng: np.random.Generator) -> np.ndarray:
    """Sinuous channels (code 1) with levee rims (code 2) in background (code 0):
    centerlines y = y0 + A sin(2 pi x / lambda + phi), widths 6-12, depth 3-5
    layers, added until about 25% channel; levees grow around channels until
    about 10% of the volume."""
    nx, ny, nz = shape
    T = np.zeros(shape, dtype=np.uint8)
    X = np.arange(nx)[:, None]
    Y = np.arange(ny)[None, :]
    while (T == 1).mean() < 0.25:
        y0 = rng.uniform(0, ny)
        A = rng.uniform(5, 25)
        lam = rng.uniform(40, 120)
        phi = rng.uniform(0, 2 * np.pi)
        w = rng.uniform(6, 12)
        depth = int(rng.integers(3, 6))
        z0 = int(rng.integers(0, nz - depth + 1))
        yc = y0 + A * np.sin(2 * np.pi * X / lam + phi)
        inside = np.abs(Y - yc) <= w / 2
        for z in range(z0, z0 + depth):
            T[:, :, z][inside] = 1
    rim = 1
    while (T == 2).mean() < 0.10 and rim < 10:
        from scipy.ndimage import binary_dilation
        grown = binary_dilation(T == 1, iterations=rim) & (T == 0)
        T[grown] = 2
        rim += 1
    return T
