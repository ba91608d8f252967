"""Pins the FP64 oracle (oracle/iqref.c + iqref.py) without the GPU.

Each function is checked against something that does not share its code:
golden vectors computed with scipy / exact rationals / brute force
(tests/golden/make_golden.py), closed forms, networkx max-flow, and the
behavioural known-answer tests of the reference (test/runtests.jl).
"""
import math
import os

import numpy as np
import pytest
from scipy.signal import correlate

import iqref

G = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))


# ------------------------------------------------------------------ A5 imfilter
def test_imfilter_golden():
    got = iqref.imfilter(G["imf_img"], G["imf_krn"])
    assert np.allclose(got, G["imf_out"], rtol=1e-13, atol=0)


def test_imfilter_vs_fft_restatement_of_gpu_path():
    # imfilter.jl:18-25: real(ifft(fft(img) .* conj(fft(padkrn)))) cropped to the valid region
    rng = np.random.default_rng(0)
    img, krn = rng.random((200, 100)), rng.random((30, 10))
    pad = np.zeros_like(img)
    pad[: krn.shape[0], : krn.shape[1]] = krn
    fft = np.real(np.fft.ifftn(np.fft.fftn(img) * np.conj(np.fft.fftn(pad))))[:171, :91]
    got = iqref.imfilter(img, krn)
    assert np.abs(got - fft).max() < 1e-2  # the tolerance of runtests.jl:152
    assert np.abs(got - fft).max() < 1e-9  # and far tighter in FP64
    assert np.allclose(got, correlate(img, krn, mode="valid"), rtol=1e-12)


# -------------------------------------------------------------- A4 fastdistance
def test_fastdistance_golden_exact():
    got = iqref.fastdistance(G["fd_ti"], G["fd_kern"], G["fd_w"])
    assert np.array_equal(got, G["fd_out"])  # dyadic inputs: every step is exact


def test_fastdistance_closed_form():
    rng = np.random.default_rng(1)
    img = rng.standard_normal((14, 11, 6))
    k = rng.standard_normal((5, 4, 3))
    w = rng.random(k.shape) < 0.5
    got = iqref.fastdistance(img, k, w)
    ref = np.zeros(got.shape)
    for c in np.ndindex(*got.shape):
        win = img[c[0]:c[0] + 5, c[1]:c[1] + 4, c[2]:c[2] + 3]
        ref[c] = (w * (win - k) ** 2).sum()
    scale = (w * (img.max() ** 2 + k ** 2)).sum()
    assert np.abs(got - ref).max() <= 1e-13 * scale
    # a dropped B2 term or a wrong sign shows up far above that bound
    assert np.abs(got - ref).max() < 1e-10


def test_fastdistance_categorical_is_exact_integer():
    rng = np.random.default_rng(2)
    img = rng.integers(0, 3, (20, 20)).astype(float)
    k = rng.integers(0, 3, (6, 6)).astype(float)
    got = iqref.fastdistance(img, k)
    ref = np.array([[((img[i:i + 6, j:j + 6] - k) ** 2).sum() for j in range(15)] for i in range(15)])
    assert np.array_equal(got, ref)


# ----------------------------------------------------------------- A10 band
def test_band_golden():
    assert np.array_equal(iqref.band(G["band_D"], 0.1), G["band_out"])


def test_band_empty_mask_selects_all_enabled():
    D = np.zeros(50)
    D[[3, 7]] = np.inf
    assert np.array_equal(iqref.band(D, 0.1), np.setdiff1d(np.arange(50), [3, 7]))


# ----------------------------------------------------------- A11 relaxation
def test_relaxation_golden():
    got = iqref.relaxation(G["rel_D"], [G["rel_A"]], 0.1)
    assert np.array_equal(got, G["rel_out"])


def _relax_sets(D, Ds, tol):
    """Set formulation: top-k prefixes of the stable (value, index) orders."""
    en = int(np.isfinite(D).sum())
    allzero = np.all(D[np.isfinite(D)] == 0)
    dbsize = en if allzero else int(math.ceil(tol * en))
    top = lambda v, k: set(np.lexsort((np.arange(v.size), v))[:k].tolist())
    over = top(D, dbsize)
    frac = 0.1 * (dbsize / en)
    while True:
        k = int(math.ceil(frac * en))
        res = set(over)
        for d in Ds:
            res &= top(d, k)
            if not res:
                break
        if res:
            return np.asarray(sorted(res))
        frac = min(frac + 0.1, 1.0)


@pytest.mark.parametrize("seed", range(6))
def test_relaxation_literal_equals_set_formulation(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 600))
    D = np.round(rng.random(n) * 12)
    Ds = [np.round(rng.random(n) * 8) for _ in range(int(rng.integers(1, 4)))]
    dis = rng.random(n) < 0.15
    D[dis] = np.inf
    for d in Ds:
        d[dis] = np.inf
    tol = [0.05, 0.1, 0.3][seed % 3]
    got = iqref.relaxation(D, Ds, tol)
    assert np.array_equal(got, _relax_sets(D, Ds, tol))
    assert len(got) > 0 and np.all(np.diff(got) > 0)


# ------------------------------------------------------------- A12 taumodel
def test_taumodel_golden():
    got = iqref.taumodel(G["tau_ev"], G["tau_D1"], [G["tau_D2"]])
    # the reference rounds each operation; the golden values round once
    assert np.allclose(got, G["tau_p"], rtol=1e-13)


def test_taumodel_single_event():
    assert iqref.taumodel([4], np.arange(10.0), [])[0] == 1.0


def test_taumodel_single_source_is_rank_weight():
    rng = np.random.default_rng(3)
    D = np.round(rng.random(100) * 7)
    ev = np.arange(100)
    from scipy.stats import rankdata
    r = rankdata(D, method="dense")
    P = (100 - r + 1) / (100 - r + 1).sum()
    assert np.allclose(iqref.taumodel(ev, D, []), P, rtol=1e-13)


def test_taumodel_tied_source_is_neutral_bitwise():
    rng = np.random.default_rng(4)
    D = np.round(rng.random(80) * 7)
    ev = np.sort(rng.choice(80, 40, replace=False))
    a = iqref.taumodel(ev, D, [])
    b = iqref.taumodel(ev, D, [np.full(80, 3.0)])
    assert np.array_equal(a, b)


# --------------------------------------------------------------- A13 sample
def test_sample_small_is_first_crossing():
    rng = np.random.default_rng(5)
    for n in (1, 2, 5, 64):
        p = rng.random(n)
        for u in rng.random(20):
            t = u * p.sum()
            cum = np.cumsum(p)
            i = iqref.sample(p, u)
            expect = int(np.searchsorted(cum, t, side="left"))
            assert i == min(expect, n - 1)


def test_sample_blocked_matches_exact_rationals_within_rounding():
    from fractions import Fraction
    rng = np.random.default_rng(6)
    p = rng.random(5000)
    tot = sum(Fraction(x) for x in p)
    for u in rng.random(30):
        i = iqref.sample(p, float(u))
        t = Fraction(float(u)) * tot
        before = sum(Fraction(x) for x in p[:i])
        slack = tot * Fraction(1, 10 ** 12)
        assert before <= t + slack and before + Fraction(p[i]) >= t - slack


# ------------------------------------------------------------- A16 graphcut
def test_graphcut_reference_convention():
    # runtests.jl:112-119
    C = iqref.graphcut(np.ones((20, 20)), np.ones((20, 20)), 0)
    assert C[:-1, :].all() and not C[-1, :].any()
    C = iqref.graphcut(np.ones((20, 20)), np.ones((20, 20)), 1)
    assert C[:, :-1].all() and not C[:, -1].any()


@pytest.mark.parametrize("k", range(4))
def test_graphcut_golden_brute_force(k):
    A, B, dim = G[f"gc{k}_A"], G[f"gc{k}_B"], int(G[f"gc{k}_dim"])
    assert np.array_equal(iqref.graphcut(A, B, dim), G[f"gc{k}_mask"])
    assert np.array_equal(iqref.graphcut_ek(A, B, dim), G[f"gc{k}_mask"])


def _nx_mask(A, B, dim):
    import networkx as nx
    from networkx.algorithms.flow import boykov_kolmogorov
    shape = A.shape
    Gr = nx.DiGraph()
    idx = list(np.ndindex(*shape))
    for d in range(A.ndim):
        for c in idx:
            if c[d] + 1 >= shape[d]:
                continue
            v = tuple(c[q] + (q == d) for q in range(A.ndim))
            w = tuple(c[q] + 2 * (q == d) for q in range(A.ndim))
            ok = w[d] < shape[d]
            Du, Dv = abs(A[c] - B[c]), abs(A[v] - B[v])
            gAu, gBu = abs(A[v] - A[c]), abs(B[v] - B[c])
            gAv = abs(A[w] - A[v]) if ok else gAu
            gBv = abs(B[w] - B[v]) if ok else gBu
            cap = (Du + Dv) / (gAu + gAv + gBu + gBv + 2.0 ** -52)
            Gr.add_edge(c, v, capacity=cap)
            Gr.add_edge(v, c, capacity=cap)
    for c in idx:
        if c[dim] == 0:
            Gr.add_edge("s", c)  # no capacity attribute = infinite
        if c[dim] == shape[dim] - 1:
            Gr.add_edge(c, "t")
    R = boykov_kolmogorov(Gr, "s", "t")
    co = {"t"}
    stack = ["t"]
    while stack:
        x = stack.pop()
        for y in R.predecessors(x):
            e = R[y][x]
            if y not in co and e["capacity"] - e["flow"] > 1e-12 * max(1.0, e["capacity"] if np.isfinite(e["capacity"]) else 1.0):
                co.add(y)
                stack.append(y)
    m = np.ones(shape, dtype=bool)
    for c in idx:
        if c in co:
            m[c] = False
    return m


@pytest.mark.parametrize("shape,dim", [((5, 30), 0), ((4, 20, 5), 0), ((20, 4, 5), 1), ((6, 12, 4), 0), ((8, 16), 0)])
def test_graphcut_vs_networkx_bk(shape, dim):
    rng = np.random.default_rng(sum(shape) + dim)
    for _ in range(3):
        A = rng.standard_normal(shape)
        B = rng.standard_normal(shape)
        assert np.array_equal(iqref.graphcut(A, B, dim), _nx_mask(A, B, dim))


def test_graphcut_float32_arithmetic_differs_from_float64_only_by_rounding():
    rng = np.random.default_rng(7)
    A = rng.standard_normal((6, 20)).astype(np.float32)
    B = rng.standard_normal((6, 20)).astype(np.float32)
    m32 = iqref.graphcut(A, B, 0, is_f32=True)
    assert np.array_equal(m32, iqref.graphcut_ek(A, B, 0, is_f32=True))


# -------------------------------------------------------------- A2 geometry
@pytest.mark.parametrize("ti,tile,sim,ov,exp", [
    ((200, 200), (30, 30), (400, 400), (1 / 6, 1 / 6), dict(ovlsize=(5, 5), spacing=(25, 25), ntiles=(16, 16), padsize=(405, 405), distsize=(171, 171))),
    ((100, 100, 50), (20, 20, 5), (200, 200, 50), (0.2, 0.2, 0.2), dict(ovlsize=(4, 4, 1), spacing=(16, 16, 4), ntiles=(13, 13, 13), padsize=(212, 212, 53), distsize=(81, 81, 46))),
    ((150, 150, 30), (24, 24, 6), (300, 300, 30), (0.25, 0.25, 0.25), dict(ovlsize=(6, 6, 2), spacing=(18, 18, 4), ntiles=(17, 17, 8), padsize=(312, 312, 34), distsize=(127, 127, 25))),
    ((512, 512), (48, 48), (1024, 1024), (1 / 6, 1 / 6), dict(ovlsize=(8, 8), spacing=(40, 40), ntiles=(26, 26), padsize=(1048, 1048), distsize=(465, 465))),
    ((256, 256, 64), (32, 32, 8), (384, 384, 64), (0.25, 0.25, 0.25), dict(ovlsize=(8, 8, 2), spacing=(24, 24, 6), ntiles=(16, 16, 11), padsize=(392, 392, 68), distsize=(225, 225, 57))),
])
def test_geometry_table(ti, tile, sim, ov, exp):
    g = iqref.geometry(ti, tile, sim, ov)
    for k, v in exp.items():
        assert g[k] == v, k


# ------------------------------------------- known-answer tests (runtests.jl)
def test_homogeneous_ti_gives_homogeneous_output():
    TI = np.ones((20, 20, 20))
    r = iqref.iqsim(TI, (10, 10, 10))  # runtests.jl:11-14
    assert r[0].dtype == np.float64 and np.array_equal(r[0], TI)


def test_float32_preserved():
    TI = np.random.default_rng(0).random((20, 20, 20)).astype(np.float32)
    assert iqref.iqsim(TI, (10, 10, 10))[0].dtype == np.float32  # runtests.jl:17-19


def test_categories_from_training_image():
    TI = np.random.default_rng(1).integers(1, 4, (20, 20, 20))
    r = iqref.iqsim(TI, (10, 10, 10))  # runtests.jl:23-26
    assert set(r[0].ravel().tolist()) <= set(np.unique(TI).tolist())


def test_soft_trend_ordering():
    TI = np.concatenate([np.zeros((10, 20, 1)), np.ones((10, 20, 1))], axis=0)
    trend = np.concatenate([np.zeros((20, 10, 1)), np.ones((20, 10, 1))], axis=1)
    r = iqref.iqsim(TI, (10, 10, 1), soft=[(trend, TI)], tol=1)  # runtests.jl:31-34
    assert r[0][:, :10, :].mean() <= r[0][:, 10:, :].mean()


def test_soft_data_beyond_the_grid_is_not_read():
    # AUX is padded to padsize and read through views at grid coordinates
    # (utils.jl:74-80, iqsim.jl:222-227): voxels past padsize must not matter
    rng = np.random.default_rng(21)
    T = rng.random((30, 30))
    auxti = rng.random((30, 30))
    aux = rng.random((70, 55))
    sim = (30, 30)
    pad = iqref.geometry(T.shape, (10, 10), sim, (1 / 6, 1 / 6))["padsize"]
    assert pad[0] < 70 and pad[1] < 55
    other = aux.copy()
    other[pad[0]:, :] = 1e3
    other[:, pad[1]:] = -1e3
    u = rng.random((1, 16))
    a = iqref.iqsim(T, (10, 10), sim, soft=[(aux, auxti)], uniforms=u)
    b = iqref.iqsim(T, (10, 10), sim, soft=[(other, auxti)], uniforms=u)
    c = iqref.iqsim(T, (10, 10), sim, soft=[(aux[:pad[0], :pad[1]], auxti)], uniforms=u)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[0], c[0])


def test_hard_data_honoured():
    TI = np.ones((20, 20, 20))
    obs = np.zeros((20, 20, 20))
    data = {c: 0.0 for c in np.ndindex(20, 20, 20)}
    assert np.array_equal(iqref.iqsim(TI, (10, 10, 10), hard=data)[0], obs)  # runtests.jl:52-56
    r = iqref.iqsim(TI, (10, 10, 10), hard={(19, 19, 19): 10}, nreal=3)  # runtests.jl:59-64
    assert all(x[19, 19, 19] == 10 for x in r)


def test_masked_sphere():
    TI = np.ones((20, 20, 20))
    shape, active = {}, np.ones((20, 20, 20), dtype=bool)
    for c in np.ndindex(20, 20, 20):
        if (c[0] - 9) ** 2 + (c[1] - 9) ** 2 + (c[2] - 9) ** 2 < 25:
            shape[c] = float("nan")
            active[c] = False
    r = iqref.iqsim(TI, (10, 10, 10), hard=shape)[0]  # runtests.jl:69-80
    assert np.isnan(r[~active]).all() and not np.isnan(r[active]).any()


def test_masked_training_image():
    TI = np.ones((20, 20, 20))
    TI[:, 4, :] = np.nan
    assert np.array_equal(iqref.iqsim(TI, (10, 10, 10))[0], np.ones((20, 20, 20)))  # runtests.jl:83-86
    TI[0, 4, :] = 0
    assert np.array_equal(iqref.iqsim(TI, (10, 10, 10))[0], np.ones((20, 20, 20)))  # runtests.jl:87-89


def test_voxel_reuse_range():
    TI = np.ones((20, 20, 20))
    _, _, voxs = iqref.iqsim(TI, (10, 10, 10), overlap=(1 / 3, 1 / 3, 1 / 3), debug=True)  # runtests.jl:109-110
    assert 0 <= voxs[0] <= 1


def test_genpath_lengths_and_data_first():
    for kind in ("raster", "dilation", "random"):
        assert len(iqref.genpath(np.random.default_rng(123), (10, 10, 10), kind, [])) == 1000  # runtests.jl:124-128
    p = iqref.genpath(np.random.default_rng(123), (10, 10, 10), "raster", [0, 999])  # runtests.jl:131-133
    assert p[:2] in ([0, 999], [999, 0]) and sorted(p) == list(range(1000))


# ------------------------ composition of the tile loop (iqsim.jl:172-285)
# Per-step checks of a logged oracle simulation, derived from the reference's
# rules and not from the oracle's code: the overlap mask and the cut slabs
# follow from which neighbours were pasted before the step (iqsim.jl:188-205,
# 252-275); a cut keeps its first slice along the cut axis and hands its last
# slice to the training image (source/sink terminals, graphcut.jl:57-81), and
# the next-side slabs are negated (iqsim.jl:273); outside every slab the
# chosen window is pasted (iqsim.jl:278); the pick lies in the tolerance band
# of distances recomputed here (iqsim.jl:237).
def _sliding_sq_dist(TI, K, M, disabled):
    """D[c] = sum over mask voxels of (TI[c + v] - K[v])^2, FP64 numpy."""
    ts = K.shape
    ds = tuple(a - b + 1 for a, b in zip(TI.shape, ts))
    D = np.zeros(ds)
    for v in zip(*np.nonzero(M)):
        win = TI[tuple(slice(o, o + n) for o, n in zip(v, ds))]
        D += (win - K[v]) ** 2
    D[disabled] = np.inf
    return D


def _composition_violations(T, tr, tol, hardtiles=()):
    geo = tr["geo"]
    TI = iqref.prepare(T).astype(np.float64)
    N = TI.ndim
    ts, ov, sp, nt, pad = geo["tilesize"], geo["ovlsize"], geo["spacing"], geo["ntiles"], geo["padsize"]
    ds = geo["distsize"]
    disabled = iqref.finddisabled(T, geo)
    errs = []
    grid = np.zeros(pad)
    pasted = set()
    for s, ind in enumerate(tr["path"]):
        c = int(tr["chosen"][0, s])
        if c < 0:
            continue
        t3 = np.unravel_index(ind, nt, order="F")
        org = [t3[d] * sp[d] for d in range(N)]
        region = tuple(slice(o, o + n) for o, n in zip(org, ts))
        before, cut = tr["before"][0, s], tr["cutmask"][0, s]
        if not np.array_equal(before, grid[region]):
            errs.append((s, "grid before the step"))
        slabs = []
        for d in range(N):
            if ov[d] <= 1:
                continue
            for side, step in ((0, -1), (1, 1)):
                nb = list(t3)
                nb[d] += step
                if 0 <= nb[d] < nt[d] and np.ravel_multi_index(nb, nt, order="F") in pasted:
                    lo, hi = (0, ov[d]) if side == 0 else (sp[d], ts[d])
                    sl = np.zeros(ts, dtype=bool)
                    sl[(slice(None),) * d + (slice(lo, hi),)] = True
                    slabs.append((d, side, lo, hi, sl))
        union = np.zeros(ts, dtype=bool)
        for *_, sl in slabs:
            union |= sl
        if cut[~union].any():
            errs.append((s, "cut outside the slabs"))
        for k, (d, side, lo, hi, sl) in enumerate(slabs):
            others = np.zeros(ts, dtype=bool)
            for j, (*_, o) in enumerate(slabs):
                if j != k:
                    others |= o
            first = (slice(None),) * d + (lo,)
            last = (slice(None),) * d + (hi - 1,)
            keep_at, take_at = (first, last) if side == 0 else (last, first)
            if not cut[keep_at].all():
                errs.append((s, f"slab axis {d} side {side}: grid not kept at its grid-side face"))
            if cut[take_at][~others[take_at]].any():
                errs.append((s, f"slab axis {d} side {side}: training image not taken at its far face"))
        cs = np.unravel_index(c, ds, order="F")
        W = TI[tuple(slice(o, o + n) for o, n in zip(cs, ts))]
        if ind not in hardtiles:
            D = _sliding_sq_dist(TI, before, union, disabled)
            dmin = D.min()
            if not D.ravel(order="F")[c] <= (1 + tol) * dmin * (1 + 1e-9) + 1e-12:
                errs.append((s, "pick outside the tolerance band"))
        grid[region] = np.where(cut, before, W)
        pasted.add(ind)
    return errs, grid


def _logged(T, tile, sim, seed=0, **kw):
    return iqref.iqsim(T, tile, sim, seed=seed, return_trace=True, step_log=True, **kw)


_COMPOSITION_CASES = [
    ("random 3-D", dict(tile=(8, 8, 4), sim=(24, 24, 10), shape=(26, 24, 10), path="random", ov=(0.25, 0.25, 0.5))),
    ("dilation 2-D", dict(tile=(10, 10), sim=(50, 45), shape=(40, 40), path="dilation", ov=(0.3, 0.3))),
    ("data-first 2-D", dict(tile=(10, 10), sim=(50, 50), shape=(40, 40), path="raster", ov=(0.3, 0.3), hard=True)),
]


def _composition_case(name, mutate=0):
    cfg = dict(_COMPOSITION_CASES)[name]
    rng = np.random.default_rng(50)
    x = rng.standard_normal(cfg["shape"])
    for ax in range(x.ndim):
        x = np.apply_along_axis(lambda v: np.convolve(v, np.ones(3) / 3, mode="same"), ax, x)
    hard, hardtiles = {}, set()
    if cfg.get("hard"):
        hard = {(int(i), int(j)): float(x[i % cfg["shape"][0], j % cfg["shape"][1]])
                for i, j in zip(rng.integers(0, cfg["sim"][0], 4), rng.integers(0, cfg["sim"][1], 4))}
        geo = iqref.geometry(x.shape, cfg["tile"], cfg["sim"], cfg["ov"])
        hardtiles = set(iqref.findskipped(hard, geo)[1])
    res, tr = _logged(x, cfg["tile"], cfg["sim"], overlap=cfg["ov"], path=cfg["path"], tol=0.1, hard=hard,
                      seed=3, debug=True, mutate=mutate)
    return x, res, tr, hardtiles, hard


@pytest.mark.parametrize("name", [n for n, _ in _COMPOSITION_CASES])
def test_tile_loop_composition(name):
    x, (reals, cuts, _), tr, hardtiles, hard = _composition_case(name)
    errs, grid = _composition_violations(x, tr, 0.1, hardtiles)
    assert not errs, errs[:5]
    # the replayed pastes give the oracle's output (hard data overwrites
    # after the loop, iqsim.jl:291-294)
    geo = tr["geo"]
    crop = tuple(slice(0, n) for n in geo["simsize"])
    out = np.asarray(reals[0], dtype=np.float64)
    keep = np.ones(out.shape, dtype=bool)
    for k in hard:
        keep[k] = False
    assert np.array_equal(grid[crop][keep], out[keep])
    # next-side slabs occur, so the negation is exercised
    nexts = 0
    nt, pasted = geo["ntiles"], set()
    for s_, ind in enumerate(tr["path"]):
        t3 = np.unravel_index(ind, nt, order="F")
        for d in range(len(nt)):
            nb = list(t3)
            nb[d] += 1
            if geo["ovlsize"][d] > 1 and nb[d] < nt[d] and np.ravel_multi_index(nb, nt, order="F") in pasted:
                nexts += 1
        pasted.add(ind)
    assert nexts > 0


@pytest.mark.parametrize("mutate", [1, 2])
@pytest.mark.parametrize("name", ["random 3-D", "dilation 2-D"])
def test_composition_checks_catch_planted_mistakes(name, mutate):
    # a next slab left un-negated, or a prev slab one voxel off, must be caught
    x, _, tr, hardtiles, _ = _composition_case(name, mutate=mutate)
    errs, _ = _composition_violations(x, tr, 0.1, hardtiles)
    assert errs


def test_two_tile_golden_case():
    # Derived by hand.  Training image: one row [0 1 2 3 10 11 12 13] (8 x 1);
    # tile 4 x 1, overlap 1/2: two voxels along x, none along y; grid 4 x 1
    # -> two tiles, at x = 0..3 (tile 0) and x = 2..5 (tile 1); padsize 6.
    # Path: tile 1 first, then tile 0.
    T = np.array([[0.0], [1.0], [2.0], [3.0], [10.0], [11.0], [12.0], [13.0]])
    kw = dict(overlap=(0.5, 0.5), tol=0.1, simpath=[1, 0])
    # Step 1 (tile 1, nothing pasted): empty mask, D == 0, every one of the
    # five candidates is in the band; the forced pick 0 pastes [0 1 2 3] at
    # x = 2..5.
    # Step 2 (tile 0, its next neighbour pasted): the overlap is the next slab
    # x = 2..3 holding [0 1]; D[c] = T[c+2]^2 + (T[c+3] - 1)^2.
    D = [T[c + 2, 0] ** 2 + (T[c + 3, 0] - 1.0) ** 2 for c in range(5)]
    assert D == [8.0, 90.0, 200.0, 242.0, 288.0]
    # min 8, threshold 1.1 * 8: the band is {0}; the window is [0 1 2 3].
    # The slab's cut (two slices, no interior node): its first slice x = 2
    # is the source side, kept by the cut and negated for a next slab
    # (iqsim.jl:273), so it takes the window's 2; its last slice x = 3 keeps
    # the grid's 1.  Outside the slab the window is pasted: x = 0, 1 -> 0, 1.
    (reals, cuts, _), tr = iqref.iqsim(T, (4, 1), (4, 1), forced=[[0, 0]], debug=True, return_trace=True,
                                       step_log=True, **kw)
    assert list(tr["chosen"][0]) == [0, 0]
    assert tr["before"][0, 1][:, 0].tolist() == [0.0, 0.0, 0.0, 1.0]
    assert tr["cutmask"][0, 1][:, 0].tolist() == [False, False, False, True]
    assert reals[0][:, 0].tolist() == [0.0, 1.0, 2.0, 1.0]
    assert cuts[0][:, 0].tolist() == [0.0, 0.0, 0.0, 1.0]
    # with the next slab left un-negated the faces swap: [0 1 0 3]
    bad = iqref.iqsim(T, (4, 1), (4, 1), forced=[[0, 0]], mutate=1, **kw)
    assert bad[0][:, 0].tolist() == [0.0, 1.0, 0.0, 3.0]


# ------------------------------------------ every candidate disabled
def test_band_of_all_infinite_distances_is_everything():
    # iqsim.jl:237 with minimum(D) = Inf: D .<= Inf holds for every entry
    assert list(iqref.band(np.full(7, np.inf), 0.1)) == list(range(7))


def test_relaxation_of_all_infinite_distances_is_refused():
    # relaxation.jl:7-36 never terminates when no distance is finite
    with pytest.raises(RuntimeError, match="-3"):
        iqref.relaxation(np.full(9, np.inf), [np.full(9, np.inf)], 0.1)


# ------------------------------------------------ learned encoder (no reference counterpart)
def test_learnref_gradient_matches_central_differences():
    """The conv encoder oracle's analytic gradient (every parameter group)
    against central differences of its own loss, FP64."""
    import learnref as L

    rng = np.random.default_rng(0)
    x = rng.standard_normal((7, 6, 5))
    toff = np.array([(i, j, k) for k in range(2) for j in range(3) for i in range(3) if i < 1 or j < 1])
    V = len(toff)
    th = rng.standard_normal(L.nparams(V)) * 0.3
    nq, ns, dist = 2, 5, (5, 4, 4)
    cands = np.stack([np.stack([rng.integers(0, d, ns) for d in dist], -1) for _ in range(nq)])
    Q = rng.standard_normal((nq, V))
    D = rng.random((nq, ns)) * 3
    A2 = rng.random((nq, ns)) * 2
    tau = np.array([0.7, 1.3])
    loss, g, _ = L.loss_and_grad(x, th, toff, A2, Q, cands, D, tau)
    assert abs(loss - L.loss_only(x, th, toff, A2, Q, cands, D, tau)) < 1e-12 * abs(loss)
    C = L.C
    starts = [0, C * 27, C * 27 + C, C * 27 + C + C * C * 27, C * 27 + 2 * C + C * C * 27,
              L.NCONV - 1, L.NCONV]
    idx = np.concatenate([rng.choice(L.nparams(V), 40, replace=False), starts, [L.nparams(V) - 1]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(th)
        e[i] = h
        fd = (L.loss_only(x, th + e, toff, A2, Q, cands, D, tau) - L.loss_only(x, th - e, toff, A2, Q, cands, D, tau)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * np.abs(g).max(), i


def test_learnref_conv_adjoints():
    """<g, conv(a, w)> = <conv_adjoint_input(g, w), a> = <conv_adjoint_weight(g, a), w>,
    and a single centre tap is the identity map."""
    import learnref as L

    rng = np.random.default_rng(1)
    a = rng.standard_normal((3, 5, 4, 3))
    w = rng.standard_normal((2, 3, 27))
    g = rng.standard_normal((2, 5, 4, 3))
    lhs = (g * L.conv(a, w)).sum()
    assert abs(lhs - (L.conv_adjoint_input(g, w) * a).sum()) < 1e-10 * abs(lhs)
    assert abs(lhs - (L.conv_adjoint_weight(g, a) * w).sum()) < 1e-10 * abs(lhs)
    eye = np.zeros((3, 3, 27))
    for c in range(3):
        eye[c, c, 13] = 1.0
    assert np.array_equal(L.conv(a, eye), a)
    # an x shift by one voxel, zero at the border
    sh = np.zeros((1, 1, 27))
    sh[0, 0, 14] = 1.0  # d = (+1, 0, 0)
    out = L.conv(a[:1], sh)[0]
    assert np.array_equal(out[:-1], a[0, 1:]) and not out[-1].any()
