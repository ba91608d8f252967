"""FP64 CPU oracle for the ImageQuilting.jl ``iqsim`` hot path.

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of ``bench.py`` may import this module; the product package
never does.  The arithmetic lives in ``iqref.c`` (built by the Makefile next
to this file); this module restates the host-side preparation of the
reference in numpy and calls the C restatement through ctypes.

Conventions: arrays are indexed like the Julia arrays they stand for
(``a[i, j, k]`` is Julia's ``a[i+1, j+1, k+1]``) and are handed to C in
Fortran (column-major) order; indices are 0-based; hard data is a dict that
maps 0-based coordinate tuples to values.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ORA_OK = 0
ORA_ERR_NOCAND = -3
ORA_ERR_NOTINDB = -4
ORA_ERR_ITER = -5

GR_FREQ = 0  # global-relabel period of the push-relabel schedule; 0: 8 rounds, 24 above 1,024 interior nodes (the HIP kernel's rule)

_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)


def build() -> str:
    """Compile iqref.c into libiqref.so (idempotent)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "libiqref.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libiqref.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.ora_band.restype = ctypes.c_int64
        L.ora_relaxation.restype = ctypes.c_int64
        L.ora_sample.restype = ctypes.c_int64
        L.ora_graphcut.restype = ctypes.c_int
        L.ora_graphcut_ek.restype = ctypes.c_int
        L.ora_simulate.restype = ctypes.c_int
        L.ora_num_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def set_threads(n: int) -> None:
    lib().ora_set_threads(int(n))


def num_threads() -> int:
    return int(lib().ora_num_threads())


# ---------------------------------------------------------------------------
# ctypes helpers
# ---------------------------------------------------------------------------
def _isnan_value(v) -> bool:
    """isnan of one hard-data value (utils.jl:137,150) of any floating type."""
    return isinstance(v, (float, np.floating)) and bool(np.isnan(v))


def _d3(shape: Sequence[int]) -> np.ndarray:
    s = list(shape) + [1] * (3 - len(shape))
    return np.asarray(s, dtype=np.int64)


def _f(a, dtype=np.float64) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


# ---------------------------------------------------------------------------
# A5 / A4: correlation and masked distance (imfilter.jl:5-26, utils.jl:5-13)
# ---------------------------------------------------------------------------
def imfilter(img, krn) -> np.ndarray:
    img = _f(img)
    krn = _f(krn)
    idims, kdims = _d3(img.shape), _d3(krn.shape)
    oshape = tuple(int(a - b + 1) for a, b in zip(img.shape, krn.shape))
    out = np.zeros(oshape, dtype=np.float64, order="F")
    lib().ora_imfilter(_p(img, _f64p), _p(idims, _i64p), _p(krn, _f64p), _p(kdims, _i64p), _p(out, _f64p))
    return out


def fastdistance(img, kern, weights=None) -> np.ndarray:
    img = _f(img)
    kern = _f(kern)
    idims, kdims = _d3(img.shape), _d3(kern.shape)
    w = None if weights is None else _f(weights, np.uint8)
    oshape = tuple(int(a - b + 1) for a, b in zip(img.shape, kern.shape))
    out = np.zeros(oshape, dtype=np.float64, order="F")
    lib().ora_fastdistance(_p(img, _f64p), _p(idims, _i64p), _p(kern, _f64p),
                           None if w is None else _p(w, _u8p), _p(kdims, _i64p), _p(out, _f64p))
    return out


# ---------------------------------------------------------------------------
# A10-A13: selection (iqsim.jl:237-243, relaxation.jl, taumodel.jl)
# ---------------------------------------------------------------------------
def band(D, tol: float) -> np.ndarray:
    D = np.ascontiguousarray(np.ravel(D, order="F"), dtype=np.float64)
    out = np.zeros(D.size, dtype=np.int64)
    k = lib().ora_band(_p(D, _f64p), ctypes.c_int64(D.size), ctypes.c_double(tol), _p(out, _i64p))
    return out[:k].copy()


def _maps(Ds):
    arrs = [np.ascontiguousarray(np.ravel(d, order="F"), dtype=np.float64) for d in Ds]
    ptrs = (ctypes.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
    return arrs, ptrs


def relaxation(D, Ds, tol: float) -> np.ndarray:
    D = np.ascontiguousarray(np.ravel(D, order="F"), dtype=np.float64)
    arrs, ptrs = _maps(Ds)
    out = np.zeros(D.size, dtype=np.int64)
    k = lib().ora_relaxation(_p(D, _f64p), ptrs, ctypes.c_int(len(arrs)), ctypes.c_int64(D.size),
                             ctypes.c_double(tol), _p(out, _i64p))
    if k < 0:
        raise RuntimeError(f"relaxation: error {k}")
    return out[:k].copy()


def taumodel(events, D1, Ds) -> np.ndarray:
    ev = np.ascontiguousarray(events, dtype=np.int64)
    D1 = np.ascontiguousarray(np.ravel(D1, order="F"), dtype=np.float64)
    arrs, ptrs = _maps(Ds)
    p = np.zeros(max(1, ev.size), dtype=np.float64)
    lib().ora_taumodel(_p(ev, _i64p), ctypes.c_int64(ev.size), _p(D1, _f64p), ptrs,
                       ctypes.c_int(len(arrs)), _p(p, _f64p))
    return p[: ev.size].copy()


def sample(p, u: float) -> int:
    p = np.ascontiguousarray(p, dtype=np.float64)
    return int(lib().ora_sample(_p(p, _f64p), ctypes.c_int64(p.size), ctypes.c_double(u)))


# ---------------------------------------------------------------------------
# A16: graph cut (graphcut.jl:5-84)
# ---------------------------------------------------------------------------
def graphcut(A, B, dim: int, is_f32: Optional[bool] = None, gr_freq: int = GR_FREQ,
             return_iters: bool = False):
    """Cut mask of the slab pair (A, B) along 0-based axis ``dim``."""
    if is_f32 is None:
        is_f32 = np.asarray(A).dtype == np.float32
    shape = np.asarray(A).shape
    A = _f(A)
    B = _f(B)
    s = _d3(shape)
    m = np.zeros(shape, dtype=np.uint8, order="F")
    it = np.zeros(1, dtype=np.int64)
    rc = lib().ora_graphcut(_p(A, _f64p), _p(B, _f64p), ctypes.c_int(int(is_f32)), _p(s, _i64p),
                            ctypes.c_int(dim), ctypes.c_int(gr_freq), _p(m, _u8p), _p(it, _i64p))
    if rc != ORA_OK:
        raise RuntimeError(f"graphcut: error {rc}")
    m = m.astype(bool)
    return (m, int(it[0])) if return_iters else m


def graphcut_ek(A, B, dim: int, is_f32: Optional[bool] = None) -> np.ndarray:
    """Independent Edmonds-Karp cross-check of :func:`graphcut`."""
    if is_f32 is None:
        is_f32 = np.asarray(A).dtype == np.float32
    shape = np.asarray(A).shape
    A = _f(A)
    B = _f(B)
    s = _d3(shape)
    m = np.zeros(shape, dtype=np.uint8, order="F")
    lib().ora_graphcut_ek(_p(A, _f64p), _p(B, _f64p), ctypes.c_int(int(is_f32)), _p(s, _i64p),
                          ctypes.c_int(dim), _p(m, _u8p))
    return m.astype(bool)


# ---------------------------------------------------------------------------
# Host preparation, restated from the reference (numpy)
# ---------------------------------------------------------------------------
def geometry(tisize, tilesize, simsize, overlap):
    """iqsim.jl:92-115."""
    tilesize = tuple(int(t) for t in tilesize)
    ovlsize = tuple(int(math.ceil(o * t)) for o, t in zip(overlap, tilesize))
    spacing = tuple(t - o for t, o in zip(tilesize, ovlsize))
    ntiles = tuple(int(math.ceil(s / max(sp, 1))) for s, sp in zip(simsize, spacing))
    padsize = tuple(n * (t - o) + o for n, t, o in zip(ntiles, tilesize, ovlsize))
    distsize = tuple(T - t + 1 for T, t in zip(tisize, tilesize))
    ovlvol = int(np.prod(padsize) - np.prod([p - (n - 1) * o for p, n, o in zip(padsize, ntiles, ovlsize)]))
    return dict(tilesize=tilesize, ovlsize=ovlsize, spacing=spacing, ntiles=ntiles, padsize=padsize,
                distsize=distsize, TIsize=tuple(tisize), simsize=tuple(simsize), ovlvol=ovlvol)


def _unmask(img):
    """Masked entries of a numpy masked array are Julia `missing` (None)."""
    if isinstance(img, np.ma.MaskedArray):
        out = np.asarray(img.data).astype(object)
        out[np.ma.getmaskarray(img)] = None
        return out
    return img


def prepare(img):
    """utils.jl:96-102: missing (None / masked) and NaN become 0; non-float becomes Float64."""
    a = np.asarray(_unmask(img))
    if a.dtype == object:
        vals = [np.nan if v is None else v for v in a.ravel(order="F")]
        base = np.asarray([v for v in vals if not _isnan_value(v)])
        ftype = base.dtype if base.dtype.kind == "f" else np.float64
        f = np.asarray(vals, dtype=np.float64).astype(ftype).reshape(a.shape, order="F")
    elif a.dtype.kind == "f":
        f = a.copy()
    else:
        f = a.astype(np.float64)
    f[np.isnan(f)] = 0
    return f


def finddisabled(trainimg, geo) -> np.ndarray:
    """utils.jl:115-129: candidate windows holding a NaN/missing voxel."""
    a = np.asarray(_unmask(trainimg))
    if a.dtype == object:
        nanmask = np.vectorize(lambda v: v is None or _isnan_value(v))(a)
    elif a.dtype.kind == "f":
        nanmask = np.isnan(a)
    else:
        nanmask = np.zeros(a.shape, dtype=bool)
    dis = np.zeros(geo["distsize"], dtype=bool)
    ts, ds = geo["tilesize"], geo["distsize"]
    for ind in zip(*np.nonzero(nanmask)):
        sl = tuple(slice(max(i - t + 1, 0), min(i, d - 1) + 1) for i, t, d in zip(ind, ts, ds))
        dis[sl] = True
    return dis


def _tile_coords(tileind, geo):
    start = [t * s for t, s in zip(tileind, geo["spacing"])]
    return start, [s + t for s, t in zip(start, geo["tilesize"])]


def findskipped(hard: Dict, geo):
    """utils.jl:131-156 (with activation!/indicator!, utils.jl:31-55)."""
    N = len(geo["ntiles"])
    nan_keys = {k for k, v in hard.items() if _isnan_value(v)}
    data_keys = {k for k, v in hard.items() if not _isnan_value(v)}
    skipped, datainds = set(), []
    for lin in range(int(np.prod(geo["ntiles"]))):
        tileind = np.unravel_index(lin, geo["ntiles"], order="F")
        start, finish = _tile_coords(tileind, geo)
        if any(s >= ss for s, ss in zip(start, geo["simsize"])):
            skipped.add(lin)
            continue
        ranges = [range(s, f) for s, f in zip(start, finish)]
        tvox = int(np.prod(geo["tilesize"]))
        ninact = sum(1 for k in nan_keys if all(r.start <= k[d] < r.stop for d, r in enumerate(ranges)))
        if ninact == tvox:
            skipped.add(lin)
        elif any(all(r.start <= k[d] < r.stop for d, r in enumerate(ranges)) for k in data_keys):
            datainds.append(lin)
    return skipped, datainds


def _dilate_diamond(grid: np.ndarray) -> np.ndarray:
    """ImageMorphology 0.4 ``dilate`` with its default structuring element,
    the radius-1 diamond: a voxel is set when it or one of its face
    neighbours is (written as a loop over voxels, independent of the
    product's shifted-slice form)."""
    out = grid.copy()
    shape = grid.shape
    for idx in zip(*np.nonzero(grid)):
        for ax in range(grid.ndim):
            for step in (-1, 1):
                j = list(idx)
                j[ax] += step
                if 0 <= j[ax] < shape[ax]:
                    out[tuple(j)] = True
    return out


def genpath(rng: np.random.Generator, extent, kind: str, datainds: List[int]) -> List[int]:
    """utils.jl:158-204 (0-based, column-major tile indices)."""
    n = int(np.prod(extent))
    if not datainds:
        if kind == "raster":
            return list(range(n))
        if kind == "random":
            return [int(i) for i in rng.permutation(n)]
        if kind == "dilation":
            pivot = int(rng.integers(n))
            flat = np.zeros(n, dtype=bool)
            flat[pivot] = True
            path = [pivot]
            grid = flat.reshape(extent, order="F")
            while not grid.all():
                dil = _dilate_diamond(grid)
                new = np.flatnonzero((dil & ~grid).ravel(order="F"))
                path.extend(int(i) for i in new)
                grid = dil
            return path
        raise ValueError(kind)
    order = [datainds[i] for i in rng.permutation(len(datainds))]
    flat = np.zeros(n, dtype=bool)
    path = []
    for p in order:
        flat[p] = True
        path.append(int(p))
    grid = flat.reshape(extent, order="F")
    while not grid.all():
        dil = _dilate_diamond(grid)
        new = np.flatnonzero((dil & ~grid).ravel(order="F"))
        path.extend(int(i) for i in new)
        grid = dil
    return path


def pad_symmetric(aux, padsize):
    """ImageFiltering ``Pad(:symmetric, 0, padsize - size)`` (utils.jl:74-80)."""
    a = np.asarray(aux)
    widths = [(0, max(p - s, 0)) for p, s in zip(padsize, a.shape)]
    return np.pad(a, widths, mode="symmetric")


# ---------------------------------------------------------------------------
# A1: iqsim restated end to end (iqsim.jl:50-315) with pre-drawn uniforms
# ---------------------------------------------------------------------------
def iqsim(trainimg, tilesize, simsize=None, *, overlap=None, soft=(), hard=None, tol=0.1,
          path="raster", nreal=1, debug=False, uniforms=None, forced=None, seed=0,
          gr_freq=GR_FREQ, return_trace=False, max_steps=None, simpath=None, step_log=False, mutate=0):
    """The tile loop of iqsim.jl:163-300 in FP64 C.  ``step_log``: the trace
    also holds, per realization and path step, the tile's grid values before
    the paste and its cut mask (``before``, ``cutmask``: nreal x npath x
    tilesize).  ``mutate`` plants a mistake in the loop (1: next slabs not
    negated, 2: prev slabs one voxel further in); only tests of the checks
    themselves use it."""
    T = np.asarray(trainimg)
    N = T.ndim
    if simsize is None:
        simsize = T.shape
    if overlap is None:
        overlap = tuple(1 / 6 for _ in range(N))
    hard = dict(hard or {})
    geo = geometry(T.shape, tilesize, simsize, overlap)
    TI = prepare(T)
    is_f32 = TI.dtype == np.float32
    SOFT = []
    for aux, auxti in soft:
        # the tile loop reads AUX at grid coordinates (views, iqsim.jl:222-227):
        # only its first padsize voxels per axis matter
        A = prepare(pad_symmetric(aux, geo["padsize"]))
        A = np.ascontiguousarray(A[tuple(slice(0, p) for p in geo["padsize"])])
        SOFT.append((A, prepare(auxti)))
    disabled = finddisabled(T, geo)
    skipped, datainds = findskipped(hard, geo)
    rng = np.random.default_rng(seed)
    if simpath is None:
        simpath = genpath(rng, geo["ntiles"], path, list(datainds))
    simpath = list(simpath)
    if max_steps is not None:  # bounded sample: the first tile steps only
        simpath = simpath[:max_steps]
    npath = len(simpath)
    if uniforms is None:
        uniforms = rng.random((nreal, npath))
    uniforms = np.ascontiguousarray(uniforms, dtype=np.float64).reshape(nreal, npath)

    d3 = lambda t: np.asarray(list(t) + [1] * (3 - len(t)), dtype=np.int64)
    padvox = int(np.prod(geo["padsize"]))
    out = np.zeros(nreal * padvox, dtype=np.float64)
    chosen = np.zeros(nreal * npath, dtype=np.int64)
    cut = np.zeros(nreal * padvox, dtype=np.float64) if debug else None
    cut_iters = np.zeros(1, dtype=np.int64)
    ti64 = _f(TI)
    dis = _f(disabled, np.uint8)
    pth = np.ascontiguousarray(simpath, dtype=np.int64)
    skp = np.zeros(int(np.prod(geo["ntiles"])), dtype=np.uint8)
    for s in skipped:
        skp[s] = 1
    hardgrid = None
    if hard:
        hardgrid = np.full(geo["padsize"], np.nan, dtype=np.float64, order="F")
        for k, v in hard.items():
            hardgrid[tuple(k)] = v
    aux_arrs = [_f(a) for a, _ in SOFT]
    auxti_arrs = [_f(b) for _, b in SOFT]
    aux_ptrs = (ctypes.c_void_p * max(1, len(SOFT)))(*[a.ctypes.data for a in aux_arrs])
    auxti_ptrs = (ctypes.c_void_p * max(1, len(SOFT)))(*[a.ctypes.data for a in auxti_arrs])
    frc = None
    if forced is not None:
        frc = np.ascontiguousarray(forced, dtype=np.int64).reshape(nreal * npath)

    class Args(ctypes.Structure):
        _fields_ = [
            ("tisize", ctypes.c_int64 * 3), ("tilesize", ctypes.c_int64 * 3), ("ovlsize", ctypes.c_int64 * 3),
            ("spacing", ctypes.c_int64 * 3), ("ntiles", ctypes.c_int64 * 3), ("padsize", ctypes.c_int64 * 3),
            ("distsize", ctypes.c_int64 * 3),
            ("ti", ctypes.c_void_p), ("ti_is_f32", ctypes.c_int), ("disabled", ctypes.c_void_p),
            ("path", ctypes.c_void_p), ("npath", ctypes.c_int64), ("skipped", ctypes.c_void_p),
            ("hardgrid", ctypes.c_void_p), ("naux", ctypes.c_int), ("aux", ctypes.c_void_p),
            ("auxti", ctypes.c_void_p), ("tol", ctypes.c_double), ("nreal", ctypes.c_int64),
            ("uniforms", ctypes.c_void_p), ("forced", ctypes.c_void_p), ("gr_freq", ctypes.c_int),
            ("out", ctypes.c_void_p), ("chosen", ctypes.c_void_p), ("cutgrid", ctypes.c_void_p),
            ("cut_iters", ctypes.c_void_p), ("nband", ctypes.c_void_p),
            ("step_cut", ctypes.c_void_p), ("step_before", ctypes.c_void_p), ("mutate", ctypes.c_int),
        ]

    a = Args()
    for name in ("tilesize", "ovlsize", "spacing", "ntiles", "padsize", "distsize"):
        getattr(a, name)[:] = list(d3(geo[name]))
    a.tisize[:] = list(d3(T.shape))
    a.ti = ti64.ctypes.data
    a.ti_is_f32 = int(is_f32)
    a.disabled = dis.ctypes.data
    a.path = pth.ctypes.data
    a.npath = npath
    a.skipped = skp.ctypes.data
    a.hardgrid = None if hardgrid is None else hardgrid.ctypes.data
    a.naux = len(SOFT)
    a.aux = ctypes.cast(aux_ptrs, ctypes.c_void_p)
    a.auxti = ctypes.cast(auxti_ptrs, ctypes.c_void_p)
    a.tol = float(tol)
    a.nreal = nreal
    a.uniforms = uniforms.ctypes.data
    a.forced = None if frc is None else frc.ctypes.data
    a.gr_freq = gr_freq
    a.out = out.ctypes.data
    a.chosen = chosen.ctypes.data
    a.cutgrid = None if cut is None else cut.ctypes.data
    a.cut_iters = cut_iters.ctypes.data
    nband = np.zeros(nreal * npath, dtype=np.int64)
    a.nband = nband.ctypes.data
    tvox = int(np.prod(geo["tilesize"]))
    step_cut = np.zeros(nreal * npath * tvox, dtype=np.uint8) if step_log else None
    step_before = np.zeros(nreal * npath * tvox, dtype=np.float64) if step_log else None
    a.step_cut = None if step_cut is None else step_cut.ctypes.data
    a.step_before = None if step_before is None else step_before.ctypes.data
    a.mutate = int(mutate)
    rc = lib().ora_simulate(ctypes.byref(a))
    if rc != ORA_OK:
        raise RuntimeError(f"oracle simulate failed: {rc}")

    reals, cuts, voxs = [], [], []
    for r in range(nreal):
        sim = out[r * padvox:(r + 1) * padvox].reshape(geo["padsize"], order="F").astype(TI.dtype)
        cg = cut[r * padvox:(r + 1) * padvox].reshape(geo["padsize"], order="F") if debug else None
        if debug:
            voxs.append(float(cg.sum()) / geo["ovlvol"])                  # iqsim.jl:288
        for k, v in hard.items():                                          # iqsim.jl:291-294
            sim[tuple(k)] = v
            if debug and _isnan_value(v):
                cg[tuple(k)] = v
        crop = tuple(slice(0, s) for s in geo["simsize"])                 # iqsim.jl:303
        sim = sim[crop]
        reals.append(unprepare(T, sim))
        if debug:
            cuts.append(cg[crop])
    trace = dict(chosen=chosen.reshape(nreal, npath), path=simpath, geo=geo, uniforms=uniforms,
                 cut_iters=int(cut_iters[0]), nband=nband.reshape(nreal, npath))
    if step_log:
        tshape = (nreal, npath) + tuple(geo["tilesize"])
        # per step a column-major tile: reversed axes in C order, then transposed
        cshape = tshape[:2] + tuple(geo["tilesize"])[::-1]
        axes = (0, 1) + tuple(range(len(tshape) - 1, 1, -1))
        trace["cutmask"] = step_cut.reshape(cshape).transpose(axes).astype(bool)
        trace["before"] = step_before.reshape(cshape).transpose(axes)
    res = (reals, cuts, voxs) if debug else reals
    return (res, trace) if return_trace else res


def unprepare(T, fimg):
    """utils.jl:104-113 / iqsim.jl:145: floats stay floats; other types map NaN to None."""
    T = np.asarray(T)
    if T.dtype.kind == "f":
        return fimg
    out = np.empty(fimg.shape, dtype=object)
    for idx in np.ndindex(fimg.shape):
        v = fimg[idx]
        out[idx] = None if math.isnan(v) else (int(v) if float(v).is_integer() else v)
    return out
