"""Python mirror of ImageQuilting.jl's API over the HIP C ABI.

``iqsim`` keeps the reference's keywords, validation, typing and return
convention (iqsim.jl:50-315); every distance map, pattern pick, boundary cut
and paste runs in libiqhip on the GPU.  Host-side work is limited to what the
reference also does on the host before the tile loop (prep.py).
"""
from __future__ import annotations

import ctypes
import math
import warnings
from typing import Dict, Optional, Sequence

import numpy as np

from . import _lib
from . import prep

_f64p = ctypes.c_void_p


def _f(a, dtype) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _d3(shape, fill: int = 1) -> np.ndarray:
    """Extents padded to three axes (missing trailing axes have extent `fill`).
    Keep the returned array referenced for as long as C reads its address."""
    return np.asarray(list(shape) + [fill] * (3 - len(shape)), dtype=np.int64)


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data


class Context:
    """One GPU and one HIP stream (iq_ctx).  Not thread safe."""

    def __init__(self, device: int = 0):
        self._L = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._L.iq_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if getattr(self, "handle", None):
            self._L.iq_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc: int):
        _lib.check(rc, self.handle)

    def pinned_empty(self, n: int, dtype) -> np.ndarray:
        """A flat array of n elements in page-locked host memory
        (iq_host_alloc), freed with the array: plan results copy into it at
        the full PCIe rate."""
        dtype = np.dtype(dtype)
        h = _Pinned(self, max(int(n) * dtype.itemsize, 1))
        raw = (ctypes.c_char * h.nbytes).from_address(h.ptr)
        raw._owner = h  # the allocation lives as long as the array's buffer
        return np.frombuffer(raw, dtype=dtype, count=int(n))


class _Pinned:
    def __init__(self, ctx: "Context", nbytes: int):
        p = ctypes.c_void_p()
        ctx.check(ctx._L.iq_host_alloc(ctx.handle, nbytes, ctypes.byref(p)))
        self.ctx, self.ptr, self.nbytes = ctx, p.value, nbytes

    def __del__(self):
        try:
            if self.ctx.handle:
                self.ctx._L.iq_host_free(self.ctx.handle, ctypes.c_void_p(self.ptr))
        except Exception:
            pass


_DEFAULT: Optional[Context] = None


def default_context() -> Context:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Context(0)
    return _DEFAULT


def device_count() -> int:
    n = ctypes.c_int(0)
    _lib.check(_lib.load().iq_device_count(ctypes.byref(n)))
    return n.value


# ---------------------------------------------------------------------------
# host preparation inside the C ABI (utils.jl:69-129): raw arrays in
# ---------------------------------------------------------------------------
_RAW_CODES = {"float32": 0, "float64": 1, "uint8": 2, "int8": 3, "uint16": 4, "int16": 5, "uint32": 6, "int32": 7,
              "uint64": 8, "int64": 9}


def _raw_array(a):
    """Column-major data, raw element-type code and missing mask (or None)
    of an input image (masked arrays and object arrays holding None carry
    missing values)."""
    T = prep.element_type(a)
    miss = prep._missing_mask(a if isinstance(a, np.ma.MaskedArray) else np.asarray(a))
    base = np.asarray(a.data) if isinstance(a, np.ma.MaskedArray) else np.asarray(a)
    if base.dtype == object:
        vals = np.zeros(base.shape, dtype=T if T.name in _RAW_CODES else np.float64)
        keep = ~miss
        vals[keep] = np.asarray(base[keep].tolist(), dtype=vals.dtype)
        base = vals
    if base.dtype == np.bool_:
        base = base.astype(np.uint8)
    if base.dtype.name not in _RAW_CODES:
        base = base.astype(np.float64)
    data = np.asfortranarray(base)
    m = np.asfortranarray(miss.astype(np.uint8)) if miss.any() else None
    return data, _RAW_CODES[data.dtype.name], m


def prepare_native(img) -> np.ndarray:
    """prepare (utils.jl:96-101) in the library: FP32 stays FP32, every other
    element type becomes FP64, missing and NaN become 0."""
    L = _lib.load()
    data, code, m = _raw_array(img)
    out = np.zeros(data.shape, dtype=np.float32 if code == 0 else np.float64, order="F")
    dims = _d3(data.shape)
    _lib.check(L.iq_prepare(data.ctypes.data, code, _ptr(m), dims.ctypes.data, out.ctypes.data))
    return out


def prepare_soft_native(aux, padsize) -> np.ndarray:
    """Soft grid as the tile loop reads it (utils.jl:74-80): symmetric
    padding to padsize, prepared, first padsize voxels per axis (FP64)."""
    L = _lib.load()
    data, code, m = _raw_array(aux)
    pad = tuple(int(p) for p in padsize)
    out = np.zeros(pad, dtype=np.float64, order="F")
    dims, pad3 = _d3(data.shape), _d3(pad)
    _lib.check(L.iq_prepare_soft(data.ctypes.data, code, _ptr(m), dims.ctypes.data, pad3.ctypes.data,
                                 out.ctypes.data))
    return out


def finddisabled_native(trainimg, tilesize) -> np.ndarray:
    """finddisabled (utils.jl:115-129) in the library."""
    L = _lib.load()
    data, code, m = _raw_array(trainimg)
    ds = tuple(s - t + 1 for s, t in zip(data.shape, tilesize))
    out = np.zeros(ds, dtype=np.uint8, order="F")
    dims, tile3 = _d3(data.shape), _d3(tilesize)
    _lib.check(L.iq_finddisabled(data.ctypes.data, code, _ptr(m), dims.ctypes.data, tile3.ctypes.data,
                                 out.ctypes.data))
    return out


# ---------------------------------------------------------------------------
# per-function entry points (each replaces one reference function)
# ---------------------------------------------------------------------------
def imfilter(img, krn, ctx: Optional[Context] = None) -> np.ndarray:
    """imfilter_cpu / imfilter_gpu (imfilter.jl:5-26): valid cross-correlation."""
    ctx = ctx or default_context()
    img = _f(img, np.float64)
    krn = _f(krn, np.float64)
    oshape = tuple(a - b + 1 for a, b in zip(img.shape, krn.shape))
    if any(s <= 0 for s in oshape):
        raise ValueError("kernel larger than image")
    out = np.zeros(oshape, dtype=np.float64, order="F")
    idims, kdims = _d3(img.shape), _d3(krn.shape)
    ctx.check(ctx._L.iq_imfilter(ctx.handle, _ptr(img), _ptr(idims), _ptr(krn), _ptr(kdims), _ptr(out)))
    return out


def fastdistance(img, kern, weights=None, ctx: Optional[Context] = None) -> np.ndarray:
    """fastdistance (utils.jl:5-13): |A2 - 2AB + B2| over the weighted voxels."""
    ctx = ctx or default_context()
    img = np.asarray(img)
    dt = _lib.IQ_F32 if img.dtype == np.float32 else _lib.IQ_F64
    imgf = _f(img, np.float32 if dt == _lib.IQ_F32 else np.float64)
    kern = _f(kern, np.float64)
    w = None if weights is None else _f(np.asarray(weights) != 0, np.uint8)
    oshape = tuple(a - b + 1 for a, b in zip(imgf.shape, kern.shape))
    out = np.zeros(oshape, dtype=np.float64, order="F")
    idims, kdims = _d3(imgf.shape), _d3(kern.shape)
    ctx.check(ctx._L.iq_fastdistance(ctx.handle, _ptr(imgf), dt, _ptr(idims), _ptr(kern), _ptr(w), _ptr(kdims),
                                     _ptr(out)))
    return out


def band_pick(D, tol: float, ctx: Optional[Context] = None) -> np.ndarray:
    """findall(vec(D .<= (1 + tol) * minimum(D))) (iqsim.jl:237), 0-based."""
    ctx = ctx or default_context()
    D = np.ascontiguousarray(np.ravel(np.asarray(D, dtype=np.float64), order="F"))
    out = np.zeros(D.size, dtype=np.int64)
    cnt = np.zeros(1, dtype=np.int64)
    ctx.check(ctx._L.iq_band_pick(ctx.handle, _ptr(D), D.size, float(tol), _ptr(out), _ptr(cnt)))
    return out[: cnt[0]].copy()


def _map_ptrs(Ds):
    arrs = [np.ascontiguousarray(np.ravel(np.asarray(d, dtype=np.float64), order="F")) for d in Ds]
    ptrs = (ctypes.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
    return arrs, ptrs


def relaxation(D, Ds, tol: float, ctx: Optional[Context] = None) -> np.ndarray:
    """relaxation (relaxation.jl:5-48), 0-based ascending pattern database."""
    ctx = ctx or default_context()
    D = np.ascontiguousarray(np.ravel(np.asarray(D, dtype=np.float64), order="F"))
    arrs, ptrs = _map_ptrs(Ds)
    out = np.zeros(D.size, dtype=np.int64)
    cnt = np.zeros(1, dtype=np.int64)
    ctx.check(ctx._L.iq_relaxation(ctx.handle, _ptr(D), ctypes.cast(ptrs, ctypes.c_void_p), len(arrs), D.size,
                                   float(tol), _ptr(out), _ptr(cnt)))
    return out[: cnt[0]].copy()


def taumodel(events, D, Ds, ctx: Optional[Context] = None) -> np.ndarray:
    """taumodel (taumodel.jl:5-45): unnormalised weights of `events`."""
    ctx = ctx or default_context()
    ev = np.ascontiguousarray(events, dtype=np.int64)
    D = np.ascontiguousarray(np.ravel(np.asarray(D, dtype=np.float64), order="F"))
    arrs, ptrs = _map_ptrs(Ds)
    p = np.zeros(ev.size, dtype=np.float64)
    ctx.check(ctx._L.iq_taumodel(ctx.handle, _ptr(ev), ev.size, _ptr(D), ctypes.cast(ptrs, ctypes.c_void_p),
                                 len(arrs), D.size, _ptr(p)))
    return p


def sample(p, u: float, ctx: Optional[Context] = None) -> int:
    """Weighted draw of one index (iqsim.jl:243) for a given uniform u."""
    ctx = ctx or default_context()
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.zeros(1, dtype=np.int64)
    ctx.check(ctx._L.iq_sample(ctx.handle, _ptr(p), p.size, float(u), _ptr(out)))
    return int(out[0])


def graphcut(A, B, dim: int, ctx: Optional[Context] = None) -> np.ndarray:
    """graphcut (graphcut.jl:5-84) along 0-based axis `dim`: True = keep A."""
    ctx = ctx or default_context()
    A = np.asarray(A)
    if A.shape != np.asarray(B).shape:
        raise ValueError("arrays must have the same size for cut")
    dt = _lib.IQ_F32 if A.dtype == np.float32 else _lib.IQ_F64
    npdt = np.float32 if dt == _lib.IQ_F32 else np.float64
    Af, Bf = _f(A, npdt), _f(B, npdt)
    m = np.zeros(A.shape, dtype=np.uint8, order="F")
    dims = _d3(A.shape)
    ctx.check(ctx._L.iq_graphcut(ctx.handle, _ptr(Af), _ptr(Bf), dt, _ptr(dims), int(dim), _ptr(m)))
    return m.astype(bool)


# ---------------------------------------------------------------------------
# iqsim (iqsim.jl:50-315)
# ---------------------------------------------------------------------------
class Plan:
    """Device-resident simulation inputs (iq_plan): TI, geometry, paths, data.

    A plan covers one or more simulation grids (``grids``: list of
    dict(padsize, npath, nreal)); realizations are numbered grid by grid and
    advance in lockstep.  Per-realization arrays (uniforms, forced picks,
    chosen picks, output grids) follow that order."""

    def __init__(self, ctx: Context, descs, keep: list, grids, dtype):
        self.ctx = ctx
        self._keep = keep
        self.grids = [dict(padsize=tuple(g["padsize"]), npath=int(g["npath"]), nreal=int(g["nreal"])) for g in grids]
        self.dtype = dtype
        self.nreal = sum(g["nreal"] for g in self.grids)
        self.real_npath = [g["npath"] for g in self.grids for _ in range(g["nreal"])]
        self.real_padsize = [g["padsize"] for g in self.grids for _ in range(g["nreal"])]
        self.npath = self.grids[0]["npath"]
        self.padsize = self.grids[0]["padsize"]
        arr = (_lib.PlanDesc * len(descs))(*descs)
        h = ctypes.c_void_p()
        ctx.check(ctx._L.iq_plan_create_multi(ctx.handle, arr, len(descs), ctypes.byref(h)))
        self.handle = h

    def _per_real(self, a, dtype):
        """Concatenate per-realization rows (realization order) into one array."""
        if a is None:
            return None
        if isinstance(a, (list, tuple)):
            flat = np.concatenate([np.asarray(x, dtype=dtype).ravel() for x in a])
        else:
            flat = np.asarray(a, dtype=dtype).ravel()
        if flat.size != sum(self.real_npath):
            raise ValueError("need one entry per path position of every realization")
        return np.ascontiguousarray(flat)

    def run(self, uniforms=None, forced=None) -> None:
        u = self._per_real(uniforms, np.float64)
        f = self._per_real(forced, np.int64)
        self.ctx.check(self.ctx._L.iq_plan_run(self.ctx.handle, self.handle, _ptr(u), _ptr(f)))

    def last_time(self) -> float:
        t = ctypes.c_double(0.0)
        self.ctx.check(self.ctx._L.iq_plan_last_time(self.handle, ctypes.byref(t)))
        return t.value

    def profile(self, enable: bool = True) -> None:
        self.ctx.check(self.ctx._L.iq_plan_profile(self.handle, int(bool(enable))))

    def scan_stats(self):
        """(launches, seconds, algorithmic flop, issued MFMA flop) of the
        distance scan during the last run (profiling enabled)."""
        n = np.zeros(1, dtype=np.int64)
        t = np.zeros(3, dtype=np.float64)
        self.ctx.check(self.ctx._L.iq_plan_scan_stats(self.handle, _ptr(n), t[0:].ctypes.data, t[1:].ctypes.data,
                                                      t[2:].ctypes.data))
        return int(n[0]), float(t[0]), float(t[1]), float(t[2])

    def scan_tuning(self):
        """Wavefronts per MFMA-scan workgroup chosen at plan creation for
        launches of one and of several 32-query tiles, and the measured
        microseconds per launch {(tiles, wavefronts): us} (diagnostic)."""
        nw = np.zeros(2, dtype=np.int32)
        us = np.zeros(4, dtype=np.float32)
        self.ctx.check(self.ctx._L.iq_plan_scan_tuning(self.handle, _ptr(nw), _ptr(us)))
        meas = {(t, w): float(us[2 * (t - 1) + (0 if w == 2 else 1)]) for t in (1, 2) for w in (2, 4)}
        return (int(nw[0]), int(nw[1])), {k: v for k, v in meas.items() if v > 0}

    def cut_time(self):
        """(launches, seconds) of the min-cut kernel during the last run
        (profiling enabled)."""
        n = np.zeros(1, dtype=np.int64)
        t = np.zeros(1, dtype=np.float64)
        self.ctx.check(self.ctx._L.iq_plan_cut_time(self.handle, _ptr(n), _ptr(t)))
        return int(n[0]), float(t[0])

    def bank_stats(self):
        """(launches, seconds, algorithmic bytes) of the learned shortlist's
        bank scan during the last run (profiling enabled)."""
        n = np.zeros(1, dtype=np.int64)
        t = np.zeros(2, dtype=np.float64)
        self.ctx.check(self.ctx._L.iq_plan_bank_stats(self.handle, _ptr(n), t[0:].ctypes.data, t[1:].ctypes.data))
        return int(n[0]), float(t[0]), float(t[1])

    def select_stats(self):
        """(plain queries, summed FP64-rechecked superset sizes, summed band sizes) of the last run."""
        v = np.zeros(3, dtype=np.int64)
        self.ctx.check(self.ctx._L.iq_plan_select_stats(self.handle, v[0:].ctypes.data, v[1:].ctypes.data,
                                                        v[2:].ctypes.data))
        return int(v[0]), int(v[1]), int(v[2])

    def cut_stats(self):
        """(push rounds, global-relabel passes, cuts) of the last run."""
        v = np.zeros(3, dtype=np.int64)
        self.ctx.check(self.ctx._L.iq_plan_cut_stats(self.handle, v[0:].ctypes.data, v[1:].ctypes.data,
                                                     v[2:].ctypes.data))
        return int(v[0]), int(v[1]), int(v[2])

    # -- learned proposal (no counterpart in the reference)
    def learn(self, nsamples: int = 8192, iters: int = 16, seed: int = 0) -> None:
        """Learn the per-mask encoders and encode every candidate (iq_plan_learn)."""
        self.ctx.check(self.ctx._L.iq_plan_learn(self.ctx.handle, self.handle, int(nsamples), int(iters), int(seed)))

    def set_proposal(self, enable: bool = True, K: int = 4096, check_every: int = 8, min_recall: float = 0.99,
                     check_queries: int = 0) -> None:
        """Shortlist the K best approximate distances instead of the full scan;
        every check_every-th batch the full scan decides and measures recall,
        on an evenly strided sample of check_queries of its queries (0: on all
        of them).  A cumulative recall below min_recall returns the run to the
        full scan once 256 queries have been checked, or at once when it is
        below 1 - 4 (1 - min_recall)."""
        self.ctx.check(self.ctx._L.iq_plan_set_proposal(self.handle, int(bool(enable)), int(K), int(check_every),
                                                        float(min_recall)))
        self.ctx.check(self.ctx._L.iq_plan_set_check_sample(self.handle, int(check_queries)))

    def proposal_stats(self) -> dict:
        v = np.zeros(5, dtype=np.int64)
        self.ctx.check(self.ctx._L.iq_plan_proposal_stats(self.handle, v.ctypes.data))
        members, hits, queries, full, fell = (int(x) for x in v)
        return dict(band_members=members, in_shortlist=hits, queries=queries, fully_recalled=full,
                    recall=hits / members if members else float("nan"), fell_back=bool(fell))

    def encoder(self, mask: int):
        """(W, tvox): encoder of one mask shape (64 x V) and the mask's
        column-major tile offsets (ascending)."""
        V = ctypes.c_int64(0)
        self.ctx.check(self.ctx._L.iq_plan_encoder(self.handle, int(mask), None, None, ctypes.byref(V), None))
        W = np.zeros((64, V.value), dtype=np.float32)
        tv = np.zeros(V.value, dtype=np.int32)
        self.ctx.check(self.ctx._L.iq_plan_encoder(self.handle, int(mask), W.ctypes.data, tv.ctypes.data, None, None))
        return W, tv

    # -- convolutional encoder (no counterpart in the reference)
    def learn_conv(self, nqueries: int = 640, topn: int = 256, nrandom: int = 64, steps: int = 200, batch: int = 16,
                   lr: float = 1e-3, lr_head: float = 0.0, seed: int = 0, val_k: int = 0, nhard: int = 64) -> dict:
        """Train the convolutional encoder on (query, exact distance) pairs of
        the last run's grids and rebuild the banks from it (iq_plan_learn_conv);
        returns its statistics."""
        o = _lib.ConvOpts(int(nqueries), int(topn), int(nrandom), int(steps), int(batch), float(lr), float(lr_head),
                          int(seed), int(val_k), int(nhard))
        self.ctx.check(self.ctx._L.iq_plan_learn_conv(self.ctx.handle, self.handle, ctypes.byref(o)))
        return self.conv_stats()

    def conv_stats(self) -> dict:
        v = np.zeros(12, dtype=np.float64)
        self.ctx.check(self.ctx._L.iq_plan_conv_stats(self.handle, v.ctypes.data))
        return dict(harvest_seconds=float(v[0]), train_seconds=float(v[1]), loss_first=float(v[2]),
                    loss_last=float(v[3]), pairs=int(v[4]), steps=int(v[5]), queries=int(v[6]), kept_step=int(v[7]),
                    heldout_loss_start=float(v[8]), heldout_loss_kept=float(v[9]),
                    heldout_recall_start=float(v[10]), heldout_recall_kept=float(v[11]))

    def conv_params(self) -> np.ndarray:
        """The network's parameters (IQ_CONV_NPARAM floats)."""
        th = np.zeros(_lib.IQ_CONV_NPARAM, dtype=np.float32)
        self.ctx.check(self.ctx._L.iq_plan_conv_params(self.handle, th.ctypes.data))
        return th

    def conv_grad(self, mask: int, theta, Q, cands, D, tau):
        """(loss, gradient) of the ranking loss for parameters theta (network,
        then the mask's head) on the given pairs; the plan is not changed."""
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        Q = np.ascontiguousarray(Q, dtype=np.float32)
        cands = np.ascontiguousarray(cands, dtype=np.int32)
        D = np.ascontiguousarray(D, dtype=np.float64)
        tau = np.ascontiguousarray(tau, dtype=np.float64)
        nq, ns = cands.shape
        loss = ctypes.c_double(0.0)
        grad = np.zeros_like(theta)
        self.ctx.check(self.ctx._L.iq_plan_conv_grad(self.ctx.handle, self.handle, int(mask), theta.ctypes.data, nq, ns,
                                                     Q.ctypes.data, cands.ctypes.data, D.ctypes.data, tau.ctypes.data,
                                                     ctypes.byref(loss), grad.ctypes.data))
        return loss.value, grad

    def bank_rows(self, mask: int, cands) -> np.ndarray:
        """Bank rows (n x 64, bf16 values as float) of the given candidates."""
        cands = np.ascontiguousarray(cands, dtype=np.int32)
        out = np.zeros((cands.size, 64), dtype=np.float32)
        self.ctx.check(self.ctx._L.iq_plan_bank_rows(self.ctx.handle, self.handle, int(mask), cands.ctypes.data,
                                                     cands.size, out.ctypes.data))
        return out

    def mask_voxels(self, mask: int) -> np.ndarray:
        """Column-major tile offsets (ascending) of one overlap-mask shape."""
        V = ctypes.c_int64(0)
        self.ctx.check(self.ctx._L.iq_plan_encoder(self.handle, int(mask), None, None, ctypes.byref(V), None))
        tv = np.zeros(V.value, dtype=np.int32)
        self.ctx.check(self.ctx._L.iq_plan_encoder(self.handle, int(mask), None, tv.ctypes.data, None, None))
        return tv

    def nmasks(self) -> int:
        nm = ctypes.c_int(0)
        self.ctx._L.iq_plan_encoder(self.handle, 0, None, None, None, ctypes.byref(nm))
        return nm.value

    def results(self, grids=True, chosen=True, cuts=False, out=None):
        """Per realization: padded grids, chosen picks (npath), cut grids.
        ``out``: a flat array of the plan's element type holding every
        realization's padded grid, reused instead of a new one."""
        vox = [int(np.prod(p)) for p in self.real_padsize]
        nvox, nlog = sum(vox), sum(self.real_npath)
        if grids and out is not None:
            if out.dtype != self.dtype or out.size < max(nvox, 1) or not out.flags.c_contiguous:
                raise ValueError("out must be a contiguous array of the plan's element type and grid size")
        else:
            out = np.zeros(max(nvox, 1), dtype=self.dtype) if grids else None
        ch = np.zeros(max(nlog, 1), dtype=np.int64) if chosen else None
        cg = np.zeros(max(nvox, 1), dtype=np.float64) if cuts else None
        self.ctx.check(self.ctx._L.iq_plan_results(self.ctx.handle, self.handle, _ptr(out), _ptr(ch), _ptr(cg)))
        vo = np.concatenate([[0], np.cumsum(vox)])
        lo = np.concatenate([[0], np.cumsum(self.real_npath)])
        res = {}
        if grids:
            res["grids"] = [out[vo[r]:vo[r + 1]].reshape(self.real_padsize[r], order="F") for r in range(self.nreal)]
        if chosen:
            rows = [ch[lo[r]:lo[r + 1]] for r in range(self.nreal)]
            res["chosen"] = np.stack(rows) if len(set(self.real_npath)) == 1 else rows
        if cuts:
            res["cuts"] = [cg[vo[r]:vo[r + 1]].reshape(self.real_padsize[r], order="F") for r in range(self.nreal)]
        return res

    def close(self):
        if getattr(self, "handle", None):
            self.ctx._L.iq_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _validate(trainimg, tilesize, simsize, overlap, soft, hard, tol, path, nreal):
    """The @asserts of iqsim.jl:69-89."""
    tis = np.shape(trainimg)
    N = len(tis)
    if len(tilesize) != N or len(simsize) != N or len(overlap) != N:
        raise ValueError("tilesize, simsize and overlap need one entry per axis")
    if not all(0 < t <= s for t, s in zip(tilesize, tis)):
        raise ValueError("invalid tile size")
    if not all(s >= t for s, t in zip(simsize, tilesize)):
        raise ValueError("invalid grid size")
    if not all(0 < o < 1 for o in overlap):
        raise ValueError("overlaps must be in range (0,1)")
    if not (0 < tol <= 1):
        raise ValueError("tolerance must be in range (0,1]")
    if path not in ("raster", "dilation", "random"):
        raise ValueError("invalid simulation path")
    if nreal <= 0:
        raise ValueError("invalid number of realizations")
    for aux, auxti in soft:
        if not all(a >= s for a, s in zip(np.shape(aux), simsize)):
            raise ValueError("soft data size < grid size")
        if np.shape(auxti) != tis:
            raise ValueError("auxiliary TI must have the same size as TI")
    if hard:
        coords = np.asarray([list(k) for k in hard.keys()])
        if (coords.max(axis=0) >= np.asarray(simsize)).any():
            raise ValueError("hard data coordinates outside of grid")
        if (coords.min(axis=0) < 0).any():
            raise ValueError("hard data coordinates must be non-negative indices")


def make_plan(trainimg, tilesize, simsize=None, *, overlap=None, soft=(), hard=None, tol=0.1, path="raster",
              nreal=1, rng=None, ctx: Optional[Context] = None, simpath=None):
    """Host preparation of iqsim.jl:65-139 and the upload of everything the
    tile loop reads.  ``simpath``: tile order to use instead of genpath's
    (0-based column-major tile indices; may visit a subset of the tiles).
    Returns (plan, info)."""
    plan, infos = make_plan_multi(trainimg, tilesize, [simsize], overlap=overlap, soft=soft,
                                  hard=None if hard is None else [hard], tol=tol, path=path, nreal=nreal, rng=rng,
                                  ctx=ctx, simpaths=None if simpath is None else [simpath])
    return plan, infos[0]


def make_plan_multi(trainimg, tilesize, simsizes, *, overlap=None, soft=(), hard=None, tol=0.1, path="raster",
                    nreal=1, rng=None, ctx: Optional[Context] = None, simpaths=None):
    """One plan over several simulation grids (realizations of different
    output sizes advancing in lockstep).  ``simsizes``: one size per grid
    (None = the TI size); ``hard``: None or one dict per grid; ``soft``:
    (data, dataTI) pairs shared by all grids (each padded to its grid);
    ``nreal``: realizations per grid (int or one per grid).  Paths are drawn
    from ``rng`` grid by grid unless ``simpaths`` gives one per grid (None
    entries are drawn).  Returns (plan, [info per grid])."""
    ctx = ctx or default_context()
    T = trainimg
    tis = np.shape(T)
    N = len(tis)
    overlap = tuple(1 / 6 for _ in range(N)) if overlap is None else tuple(overlap)
    soft = list(soft)
    rng = np.random.default_rng() if rng is None else rng
    tilesize = tuple(int(t) for t in tilesize)
    ng = len(simsizes)
    if ng == 0:
        raise ValueError("no simulation grid")
    nreals = [int(nreal)] * ng if np.isscalar(nreal) else [int(n) for n in nreal]
    hards = [dict() for _ in range(ng)] if hard is None else [dict(h or {}) for h in hard]
    if len(nreals) != ng or len(hards) != ng:
        raise ValueError("one nreal and one hard-data set per grid")
    TI = prepare_native(T)
    dtype = TI.dtype.type
    ti_type = prep.element_type(T)
    keep = [TI]
    descs, grids, infos = [], [], []
    disabled = None
    SOFT_TI = [_f(prepare_native(b), np.float64) for _, b in soft]
    auxti_p = (ctypes.c_void_p * max(1, len(soft)))(*[b.ctypes.data for b in SOFT_TI])
    keep += [SOFT_TI, auxti_p]
    warned = False
    for gi in range(ng):
        simsize = tuple(tis) if simsizes[gi] is None else tuple(int(s) for s in simsizes[gi])
        hd = hards[gi]
        _validate(T, tilesize, simsize, overlap, soft, hd, tol, path, nreals[gi])
        geo = prep.geometry(tis, tilesize, simsize, overlap)
        if not warned and any(t > 1 and o == 1 for t, o in zip(tilesize, geo["ovlsize"])):
            warnings.warn("Overlaps with only 1 voxel, check tilesize/overlap configuration")  # iqsim.jl:95-97
            warned = True
        if disabled is None:
            disabled = finddisabled_native(T, tilesize)
            keep.append(disabled)
        aux = [prepare_soft_native(a, geo["padsize"]) for a, _ in soft]
        skipped, datainds = prep.findskipped(hd, geo)
        if simpaths is not None and simpaths[gi] is not None:
            simpath = [int(t) for t in simpaths[gi]]
        else:
            simpath = prep.genpath(rng, geo["ntiles"], path, list(datainds))
        nt = int(np.prod(geo["ntiles"]))
        skp = np.zeros(nt, dtype=np.uint8)
        for s_ in skipped:
            skp[s_] = 1
        pth = np.ascontiguousarray(simpath, dtype=np.int64)
        hardgrid = None
        if hd:
            hardgrid = np.full(geo["padsize"], np.nan, dtype=np.float64, order="F")
            for k, v in hd.items():
                hardgrid[tuple(k)] = v
        d = _lib.PlanDesc()
        d.ti = TI.ctypes.data
        d.ti_dtype = _lib.IQ_F32 if dtype == np.float32 else _lib.IQ_F64
        d.tisize[:] = list(_d3(tis))
        for name in ("tilesize", "ovlsize", "spacing", "ntiles", "padsize"):
            # a missing trailing axis is one voxel thick with no overlap
            getattr(d, name)[:] = list(_d3(geo[name], 0 if name == "ovlsize" else 1))
        d.disabled = disabled.ctypes.data
        d.path = pth.ctypes.data
        d.npath = len(simpath)
        d.skipped = skp.ctypes.data
        d.hardgrid = None if hardgrid is None else hardgrid.ctypes.data
        d.naux = len(soft)
        aux_p = (ctypes.c_void_p * max(1, len(soft)))(*[a.ctypes.data for a in aux])
        d.aux = ctypes.cast(aux_p, ctypes.c_void_p)
        d.auxti = ctypes.cast(auxti_p, ctypes.c_void_p)
        d.tol = float(tol)
        d.nreal = nreals[gi]
        keep += [aux, pth, skp, hardgrid, aux_p]
        descs.append(d)
        grids.append(dict(padsize=geo["padsize"], npath=len(simpath), nreal=nreals[gi]))
        infos.append(dict(geo=geo, path=simpath, skipped=skipped, datainds=datainds, hard=hd, T=ti_type, rng=rng,
                          nreal=nreals[gi]))
    plan = Plan(ctx, descs, keep, grids, dtype)
    return plan, infos


def finish(info, grids, cuts=None):
    """Hard-data overwrite, crop and output typing (iqsim.jl:287-308)."""
    geo, hard, T = info["geo"], info["hard"], info["T"]
    reals, outcuts, voxs = [], [], []
    crop = tuple(slice(0, s) for s in geo["simsize"])
    for r, g in enumerate(grids):
        g = g.copy()
        cg = None
        if cuts is not None:
            cg = cuts[r].copy()
            voxs.append(float(cg.sum()) / geo["ovlvol"] if geo["ovlvol"] > 0 else float("nan"))  # iqsim.jl:288
        for k, v in hard.items():  # iqsim.jl:291-294
            g[tuple(k)] = v
            if cg is not None and prep.isnan_value(v):
                cg[tuple(k)] = v
        reals.append(prep.unprepare(T, g[crop]))
        if cg is not None:
            outcuts.append(cg[crop])
    return reals, outcuts, voxs


def iqsim_batch(trainimg, tilesize, simsizes, *, overlap=None, soft=(), hard=None, tol=0.1, path="raster", nreal=1,
                debug=False, rng=None, uniforms=None, ctx: Optional[Context] = None, return_trace=False,
                simpaths=None):
    """iqsim over several output grids at once (one plan, lockstep waves).
    Returns one result per grid, each what ``iqsim`` returns for it."""
    plan, infos = make_plan_multi(trainimg, tilesize, simsizes, overlap=overlap, soft=soft, hard=hard, tol=tol,
                                  path=path, nreal=nreal, rng=rng, ctx=ctx, simpaths=simpaths)
    try:
        if uniforms is None:
            uniforms = [infos[0]["rng"].random(n) for n in plan.real_npath]
        plan.run(uniforms)
        res = plan.results(grids=True, chosen=return_trace, cuts=debug)
    finally:
        plan.close()
    outs, traces, r0 = [], [], 0
    for info in infos:
        n = info["nreal"]
        grids = res["grids"][r0:r0 + n]
        cuts = res["cuts"][r0:r0 + n] if debug else None
        reals, cuts, voxs = finish(info, grids, cuts)
        outs.append((reals, cuts, voxs) if debug else reals)
        if return_trace:
            ch = res["chosen"][r0:r0 + n]
            traces.append(dict(chosen=np.asarray([np.asarray(c) for c in ch]), path=info["path"], geo=info["geo"],
                               uniforms=np.asarray([np.asarray(uniforms[r]) for r in range(r0, r0 + n)])
                               if isinstance(uniforms, list) else None))
        r0 += n
    return (outs, traces) if return_trace else outs


def iqsim(trainimg, tilesize, simsize=None, *, overlap=None, soft=(), hard: Optional[Dict] = None, tol=0.1,
          path="raster", nreal=1, debug=False, showprogress=False, rng=None, uniforms=None,
          ctx: Optional[Context] = None, return_trace=False, simpath=None):
    """Image quilting simulation (iqsim.jl:50-315).

    ``hard`` maps 0-based coordinate tuples to values (NaN marks inactive
    grid voxels); ``soft`` is a list of (data, dataTI) pairs.  Draws come from
    ``rng`` (a numpy Generator) unless ``uniforms`` (nreal x len(path)) is
    given.  ``simpath`` replaces genpath's tile order (it may visit a subset
    of the tiles: tiles it leaves out are not simulated).  Returns ``reals``
    or, with ``debug``, ``(reals, cuts, voxs)``.
    """
    plan, info = make_plan(trainimg, tilesize, simsize, overlap=overlap, soft=soft, hard=hard, tol=tol, path=path,
                           nreal=nreal, rng=rng, ctx=ctx, simpath=simpath)
    try:
        if uniforms is None:
            uniforms = info["rng"].random((int(nreal), plan.npath))
        plan.run(uniforms)
        res = plan.results(grids=True, chosen=return_trace, cuts=debug)
    finally:
        plan.close()
    reals, cuts, voxs = finish(info, res["grids"], res.get("cuts"))
    out = (reals, cuts, voxs) if debug else reals
    if return_trace:
        return out, dict(chosen=res["chosen"], path=info["path"], geo=info["geo"], uniforms=uniforms)
    return out


def voxelreuse(trainimg, tilesize, *, overlap=None, nreal=10, **kwargs):
    """voxelreuse (voxelreuse.jl:18-41): mean and std of the voxel reuse."""
    N = len(tilesize)
    overlap = tuple(1 / 6 for _ in range(N)) if overlap is None else tuple(overlap)
    ovlsize = [int(math.ceil(o * t)) for o, t in zip(overlap, tilesize)]
    ntiles = [2 if o > 1 else 1 for o in ovlsize]
    simsize = tuple(n * (t - o) + o for n, t, o in zip(ntiles, tilesize, ovlsize))
    _, _, voxs = iqsim(trainimg, tilesize, simsize, overlap=overlap, nreal=nreal, debug=True, **kwargs)
    mu = float(np.mean(voxs))
    sd = float(np.std(voxs, ddof=1)) if len(voxs) > 1 else float("nan")
    return mu, sd
