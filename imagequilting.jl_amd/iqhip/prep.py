"""Host-side preparation of ``iqsim`` (restated from the reference).

Geometry (iqsim.jl:92-115), type preparation (utils.jl:96-113), soft-data
padding (utils.jl:74-80), disabled candidates (utils.jl:115-129), skipped
tiles (utils.jl:131-156) and simulation paths (utils.jl:158-204).  Indices
are 0-based; arrays are indexed like the Julia arrays they stand for.
"""
from __future__ import annotations

import math
from typing import Dict, List, Sequence, Tuple

import numpy as np


def geometry(tisize: Sequence[int], tilesize: Sequence[int], simsize: Sequence[int],
             overlap: Sequence[float]) -> dict:
    """iqsim.jl:92-115 (ovlsize uses ceil of the Float64 product, as Julia)."""
    tilesize = tuple(int(t) for t in tilesize)
    ovlsize = tuple(int(math.ceil(float(o) * t)) for o, t in zip(overlap, tilesize))
    spacing = tuple(t - o for t, o in zip(tilesize, ovlsize))
    ntiles = tuple(int(math.ceil(s / max(sp, 1))) for s, sp in zip(simsize, spacing))
    padsize = tuple(n * (t - o) + o for n, t, o in zip(ntiles, tilesize, ovlsize))
    distsize = tuple(int(T) - t + 1 for T, t in zip(tisize, tilesize))
    inner = np.prod([p - (n - 1) * o for p, n, o in zip(padsize, ntiles, ovlsize)])
    ovlvol = int(np.prod(padsize) - inner)
    return dict(tilesize=tilesize, ovlsize=ovlsize, spacing=spacing, ntiles=ntiles, padsize=padsize,
                distsize=distsize, TIsize=tuple(int(t) for t in tisize), simsize=tuple(int(s) for s in simsize),
                ovlvol=ovlvol)


def isnan_value(v) -> bool:
    """isnan of one hard-data or object-array value (utils.jl:137,150): true
    for NaN of any floating type (Python float, numpy float16/32/64)."""
    return isinstance(v, (float, np.floating)) and bool(np.isnan(v))


def _missing_mask(a: np.ndarray) -> np.ndarray:
    if isinstance(a, np.ma.MaskedArray):
        m = np.ma.getmaskarray(a)
        data = a.data
        if data.dtype.kind == "f":
            m = m | np.isnan(data)
        return m
    if a.dtype == object:
        return np.vectorize(lambda v: v is None or isnan_value(v), otypes=[bool])(a)
    if a.dtype.kind == "f":
        return np.isnan(a)
    return np.zeros(a.shape, dtype=bool)


def element_type(a) -> np.dtype:
    """nonmissingtype(eltype(img)) of utils.jl:97."""
    if isinstance(a, np.ma.MaskedArray):
        return a.data.dtype
    a = np.asarray(a)
    if a.dtype == object:
        vals = [v for v in a.ravel() if v is not None and not isnan_value(v)]
        return np.asarray(vals).dtype if vals else np.dtype(np.float64)
    return a.dtype


def prepare(img) -> np.ndarray:
    """utils.jl:96-102: missing and NaN become 0; non-float becomes Float64."""
    T = element_type(img)
    F = T if T.kind == "f" else np.dtype(np.float64)
    miss = _missing_mask(img if isinstance(img, np.ma.MaskedArray) else np.asarray(img))
    if isinstance(img, np.ma.MaskedArray):
        base = np.asarray(img.data)
    else:
        base = np.asarray(img)
    if base.dtype == object:
        out = np.zeros(base.shape, dtype=F)
        keep = ~miss
        out[keep] = np.asarray(base[keep].tolist(), dtype=F)
    else:
        out = base.astype(F, copy=True)
    out[miss] = 0
    return out


def unprepare(T: np.dtype, fimg: np.ndarray):
    """utils.jl:104-113 and iqsim.jl:145: floats stay floats; other element
    types become object arrays of T with NaN mapped to None (Union{Missing,T})."""
    if T.kind == "f":
        return fimg
    nan = np.isnan(fimg)
    out = np.empty(fimg.shape, dtype=object)
    typed = np.where(nan, 0, fimg).astype(T)
    out[...] = typed
    out[nan] = None
    return out


def pad_symmetric(aux, padsize) -> np.ndarray:
    """ImageFiltering ``Pad(:symmetric, 0, padsize .- min.(padsize, size))``:
    mirror with the edge voxel repeated (utils.jl:74-80)."""
    a = np.asarray(aux)
    widths = [(0, max(int(p) - s, 0)) for p, s in zip(padsize, a.shape)]
    if all(w == (0, 0) for w in widths):
        return a.copy()
    if a.dtype == object:
        idx = np.pad(np.arange(a.size).reshape(a.shape), widths, mode="symmetric")
        return a.ravel()[idx]
    return np.pad(a, widths, mode="symmetric")


def _box_any(mask: np.ndarray, starts: Sequence[Sequence[int]], size: Sequence[int]) -> np.ndarray:
    """Counts of `mask` inside boxes of extent `size` at the given starts."""
    N = mask.ndim
    S = np.pad(mask.astype(np.int64), [(1, 0)] * N).cumsum(0)
    for d in range(1, N):
        S = S.cumsum(d)
    starts = np.asarray(starts, dtype=np.int64).reshape(-1, N)
    tot = np.zeros(len(starts), dtype=np.int64)
    for corner in np.ndindex(*([2] * N)):
        idx = tuple(starts[:, d] + (size[d] if corner[d] else 0) for d in range(N))
        sign = (-1) ** (N - sum(corner))
        tot += sign * S[idx]
    return tot


def finddisabled(trainimg, geo) -> np.ndarray:
    """utils.jl:115-129: candidate windows that contain an inactive voxel."""
    miss = _missing_mask(trainimg if isinstance(trainimg, np.ma.MaskedArray) else np.asarray(trainimg))
    ds, ts = geo["distsize"], geo["tilesize"]
    if not miss.any():
        return np.zeros(ds, dtype=bool)
    grids = np.meshgrid(*[np.arange(n) for n in ds], indexing="ij")
    starts = np.stack([g.ravel() for g in grids], axis=1)
    cnt = _box_any(miss, starts, ts)
    return (cnt > 0).reshape(ds)


def findskipped(hard: Dict, geo) -> Tuple[set, List[int]]:
    """utils.jl:131-156: tiles beyond the grid or fully inactive are skipped;
    tiles holding hard data are listed (column-major tile order)."""
    ntiles, ts, sp, sim, pad = geo["ntiles"], geo["tilesize"], geo["spacing"], geo["simsize"], geo["padsize"]
    N = len(ntiles)
    nan_g = np.zeros(pad, dtype=bool)
    dat_g = np.zeros(pad, dtype=bool)
    for k, v in hard.items():
        if isnan_value(v):
            nan_g[tuple(k)] = True
        else:
            dat_g[tuple(k)] = True
    tiles = list(np.ndindex(*ntiles[::-1]))
    tiles = [t[::-1] for t in tiles]  # column-major order
    starts = [[t[d] * sp[d] for d in range(N)] for t in tiles]
    tvox = int(np.prod(ts))
    nnan = _box_any(nan_g, starts, ts) if hard else np.zeros(len(tiles), dtype=np.int64)
    ndat = _box_any(dat_g, starts, ts) if hard else np.zeros(len(tiles), dtype=np.int64)
    skipped, datainds = set(), []
    for lin, (t, st) in enumerate(zip(tiles, starts)):
        if any(st[d] >= sim[d] for d in range(N)) or nnan[lin] == tvox:
            skipped.add(lin)
        elif ndat[lin] > 0:
            datainds.append(lin)
    return skipped, datainds


def _dilate(grid: np.ndarray) -> np.ndarray:
    """ImageMorphology 0.4 ``dilate(grid)``: max filter over its default
    structuring element, the diamond of radius 1 (each voxel and its 2N face
    neighbours, 1-connectivity)."""
    out = grid.copy()
    for ax in range(grid.ndim):
        n = grid.shape[ax]
        if n < 2:
            continue
        lo = [slice(None)] * grid.ndim
        hi = [slice(None)] * grid.ndim
        lo[ax], hi[ax] = slice(0, n - 1), slice(1, n)
        out[tuple(lo)] |= grid[tuple(hi)]
        out[tuple(hi)] |= grid[tuple(lo)]
    return out


def _grow(path: List[int], extent) -> List[int]:
    n = int(np.prod(extent))
    flat = np.zeros(n, dtype=bool)
    flat[path] = True
    grid = flat.reshape(extent, order="F")
    while not grid.all():
        dil = _dilate(grid)
        path.extend(int(i) for i in np.flatnonzero((dil & ~grid).ravel(order="F")))
        grid = dil
    return path


def genpath(rng: np.random.Generator, extent, kind: str, datainds: List[int]) -> List[int]:
    """utils.jl:158-204: raster (column-major), random, dilation or
    data-first (when hard data exists) tile order, 0-based."""
    n = int(np.prod(extent))
    if not datainds:
        if kind == "raster":
            return list(range(n))
        if kind == "random":
            return [int(i) for i in rng.permutation(n)]
        if kind == "dilation":
            return _grow([int(rng.integers(n))], extent)
        raise ValueError(f"invalid simulation path {kind!r}")
    shuffled = [int(datainds[i]) for i in rng.permutation(len(datainds))]
    return _grow(shuffled, extent)
