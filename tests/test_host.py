"""Host logic of the product (no GPU): the C ABI library loads and exports
every declared symbol, and the host preparation restated in iqhip.prep agrees
with the oracle's independent restatement and with the reference's rules."""
import os
import re
import subprocess

import numpy as np
import pytest

import iqhip
import iqref
from iqhip import prep

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "iqhip.h")).read()
    return sorted(set(re.findall(r"\b(iq_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = iqhip._lib.load()
    decl = _declared()
    assert set(decl) == set(iqhip.EXPORTS)
    for name in decl:
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", iqhip.LIB_PATH], capture_output=True, text=True).stdout
    for name in decl:
        assert re.search(rf"\bT {name}\b", out), name


def test_no_device_fails_loudly():
    # without a visible GPU the ABI reports an error instead of computing on the CPU
    L = iqhip._lib.load()
    import ctypes
    n = ctypes.c_int(-1)
    rc = L.iq_device_count(ctypes.byref(n))
    if rc == 0 and n.value > 0:
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    assert L.iq_ctx_create(0, ctypes.byref(h)) != 0
    assert L.iq_last_error(None)


@pytest.mark.parametrize("ti,tile,sim,ov", [((200, 200), (30, 30), (400, 400), (1 / 6, 1 / 6)),
                                            ((100, 100, 50), (20, 20, 5), (200, 200, 50), (0.2, 0.2, 0.2)),
                                            ((20, 20, 20), (10, 10, 10), (20, 20, 20), (1 / 3, 1 / 3, 1 / 3)),
                                            ((37, 21), (7, 9), (50, 31), (0.3, 0.4))])
def test_geometry_matches_oracle(ti, tile, sim, ov):
    a = prep.geometry(ti, tile, sim, ov)
    b = iqref.geometry(ti, tile, sim, ov)
    for k in b:
        assert a[k] == b[k], k


def test_prepare_unprepare():
    x = np.array([[1.0, np.nan], [2.0, 3.0]], dtype=np.float32)
    p = prep.prepare(x)
    assert p.dtype == np.float32 and p[0, 1] == 0
    i = np.array([[1, 2], [3, 4]])
    assert prep.prepare(i).dtype == np.float64
    o = np.array([[1, None], [3, 4]], dtype=object)
    assert np.array_equal(prep.prepare(o), [[1.0, 0.0], [3.0, 4.0]])
    u = prep.unprepare(np.dtype(np.int64), np.array([[1.0, np.nan]]))
    assert u[0, 0] == 1 and u[0, 1] is None
    f = prep.unprepare(np.dtype(np.float32), np.array([1.0], dtype=np.float32))
    assert f.dtype == np.float32


def test_finddisabled_matches_oracle():
    rng = np.random.default_rng(0)
    T = rng.random((30, 25, 6))
    T[rng.random(T.shape) < 0.002] = np.nan
    g = prep.geometry(T.shape, (7, 5, 3), (40, 40, 6), (0.3, 0.3, 0.3))
    assert np.array_equal(prep.finddisabled(T, g), iqref.finddisabled(T, g))


def test_findskipped_matches_oracle():
    rng = np.random.default_rng(1)
    g = prep.geometry((30, 30), (8, 8), (45, 45), (0.25, 0.25))
    hard = {}
    for c in zip(rng.integers(0, 45, 30), rng.integers(0, 45, 30)):
        hard[(int(c[0]), int(c[1]))] = float(rng.random())
    for x in range(8):
        for y in range(8):
            hard[(x, y)] = float("nan")  # tile 0 fully inactive
    a = prep.findskipped(hard, g)
    b = iqref.findskipped(hard, g)
    assert a[0] == b[0] and a[1] == b[1]
    assert 0 in a[0]


def _l1_ring_order(extent, seeds):
    """Tiles after `seeds` in the order repeated dilations with ImageMorphology
    0.4's default diamond structuring element add them (utils.jl:176-200):
    by Manhattan distance to the nearest seed, each ring in ascending
    column-major index order (findall over the grid, utils.jl:182)."""
    n = int(np.prod(extent))
    coords = np.stack(np.unravel_index(np.arange(n), extent, order="F"), axis=1)
    S = coords[list(seeds)]
    dist = np.abs(coords[:, None, :] - S[None, :, :]).sum(-1).min(1)
    seen = set(seeds)
    return [t for t in sorted(range(n), key=lambda t: (dist[t], t)) if t not in seen]


@pytest.mark.parametrize("impl", ["product", "oracle"])
def test_genpath_closed_forms(impl):
    gen = prep.genpath if impl == "product" else iqref.genpath
    for extent in ((6, 5, 4), (7, 3), (1, 9), (4, 4, 1)):
        n = int(np.prod(extent))
        assert gen(np.random.default_rng(5), extent, "raster", []) == list(range(n))  # utils.jl:162-166
        assert gen(np.random.default_rng(5), extent, "random", []) == \
            [int(i) for i in np.random.default_rng(5).permutation(n)]
        p = gen(np.random.default_rng(5), extent, "dilation", [])
        assert p[0] == int(np.random.default_rng(5).integers(n))
        assert p[1:] == _l1_ring_order(extent, [p[0]])
        data = [1, n - 1, n // 2]
        p = gen(np.random.default_rng(8), extent, "raster", list(data))  # data-first, utils.jl:186-201
        order = [data[i] for i in np.random.default_rng(8).permutation(len(data))]
        assert p[:3] == order and p[3:] == _l1_ring_order(extent, order)
    # runtests.jl:131-133: the data tiles come first
    p = gen(np.random.default_rng(123), (10, 10, 10), "raster", [0, 999])
    assert p[:2] in ([0, 999], [999, 0]) and len(p) == 1000


def test_dilation_rings_are_diamond_shells():
    # one dilation of a single voxel adds exactly its 2N face neighbours
    g = np.zeros((5, 5, 5), dtype=bool)
    g[2, 2, 2] = True
    d = prep._dilate(g)
    assert d.sum() == 7 and d[1, 2, 2] and d[2, 3, 2] and not d[1, 1, 2]


def test_findskipped_float32_nan_is_inactive():
    # hard values of any floating type: NaN marks an inactive voxel
    # (utils.jl:131-156 tests isnan); np.float32 NaN behaves as a float NaN
    g = prep.geometry((30, 30), (4, 4), (16, 16), (0.25, 0.25))
    for impl in (prep.findskipped, iqref.findskipped):
        want = impl({(i, j): float("nan") for i in range(4) for j in range(4)}, g)
        got = impl({(i, j): np.float32("nan") for i in range(4) for j in range(4)}, g)
        assert got == want and 0 in got[0] and got[1] == []
    assert prep.isnan_value(np.float32("nan")) and prep.isnan_value(float("nan"))
    assert not prep.isnan_value(np.int64(3)) and not prep.isnan_value(None)


def test_pad_symmetric_edge_repeats():
    # ImageFiltering Pad(:symmetric): mirror about the border, edge voxel repeated
    a = np.arange(4.0)
    assert np.array_equal(prep.pad_symmetric(a, (7,)), [0, 1, 2, 3, 3, 2, 1])
    b = np.arange(6.0).reshape(2, 3)
    assert prep.pad_symmetric(b, (3, 4)).shape == (3, 4)


def test_validation_mirrors_reference_asserts():
    T = np.zeros((10, 10))
    with pytest.raises(ValueError, match="tile size"):
        iqhip.api._validate(T, (11, 5), (10, 10), (0.2, 0.2), [], {}, 0.1, "raster", 1)
    with pytest.raises(ValueError, match="grid size"):
        iqhip.api._validate(T, (5, 5), (4, 10), (0.2, 0.2), [], {}, 0.1, "raster", 1)
    with pytest.raises(ValueError, match="overlap"):
        iqhip.api._validate(T, (5, 5), (10, 10), (0.0, 0.2), [], {}, 0.1, "raster", 1)
    with pytest.raises(ValueError, match="tolerance"):
        iqhip.api._validate(T, (5, 5), (10, 10), (0.2, 0.2), [], {}, 0.0, "raster", 1)
    with pytest.raises(ValueError, match="path"):
        iqhip.api._validate(T, (5, 5), (10, 10), (0.2, 0.2), [], {}, 0.1, "spiral", 1)
    with pytest.raises(ValueError, match="realizations"):
        iqhip.api._validate(T, (5, 5), (10, 10), (0.2, 0.2), [], {}, 0.1, "raster", 0)
    with pytest.raises(ValueError, match="outside"):
        iqhip.api._validate(T, (5, 5), (10, 10), (0.2, 0.2), [], {(10, 0): 1.0}, 0.1, "raster", 1)


# ----------------------------------------------- host preparation in the C ABI
def _raw_cases():
    rng = np.random.default_rng(30)
    f32 = rng.standard_normal((7, 6, 3)).astype(np.float32)
    f32[1, 2, 0] = np.nan
    f64 = rng.standard_normal((9, 5))
    f64[0, 0] = np.nan
    u8 = rng.integers(0, 3, (8, 8, 2)).astype(np.uint8)
    i64 = rng.integers(-5, 5, (6, 7)).astype(np.int64)
    obj = np.array([[1, None, 3], [None, 5, 6]], dtype=object)
    masked = np.ma.masked_array(rng.integers(0, 4, (5, 5)).astype(np.int32), mask=rng.random((5, 5)) < 0.2)
    boolean = rng.random((4, 6)) > 0.5
    return [f32, f64, u8, i64, obj, masked, boolean]


def test_native_prepare_matches_oracle():
    for a in _raw_cases():
        got = iqhip.api.prepare_native(a)
        ref = iqref.prepare(a)
        assert got.dtype == ref.dtype and np.array_equal(got, ref), a.dtype


@pytest.mark.parametrize("shape,pad", [((10, 7), (16, 16)), ((30, 40), (22, 16)), ((3, 4, 2), (13, 5, 9)),
                                       ((12, 12, 4), (12, 12, 4))])
def test_native_soft_padding_matches_oracle(shape, pad):
    rng = np.random.default_rng(31)
    a = rng.standard_normal(shape)
    a.flat[::7] = np.nan
    got = iqhip.api.prepare_soft_native(a, pad)
    ref = iqref.prepare(iqref.pad_symmetric(a, pad))[tuple(slice(0, p) for p in pad)]
    assert got.shape == tuple(pad) and np.array_equal(got, ref)
    u = rng.integers(0, 3, shape).astype(np.uint8)
    assert np.array_equal(iqhip.api.prepare_soft_native(u, pad),
                          iqref.prepare(iqref.pad_symmetric(u, pad))[tuple(slice(0, p) for p in pad)])


def test_native_finddisabled_matches_oracle():
    rng = np.random.default_rng(32)
    for shape, tile in (((20, 18), (5, 4)), ((12, 10, 6), (3, 4, 2)), ((9, 9), (9, 9))):
        T = rng.standard_normal(shape)
        T[rng.random(shape) < 0.02] = np.nan
        geo = iqref.geometry(shape, tile, shape, tuple(0.3 for _ in shape))
        assert np.array_equal(iqhip.api.finddisabled_native(T, tile).astype(bool), iqref.finddisabled(T, geo))
    obj = np.array([[1, None, 3, 4], [5, 6, 7, 8], [1, 2, 3, 4]], dtype=object)
    geo = iqref.geometry(obj.shape, (2, 2), obj.shape, (0.5, 0.5))
    assert np.array_equal(iqhip.api.finddisabled_native(obj, (2, 2)).astype(bool), iqref.finddisabled(obj, geo))


def _binding_pad_functions():
    """The padding function the Julia binding applies to each extent of the
    plan descriptor (read from the ccall site of IQHip.jl's iqsim_hip)."""
    src = open(os.path.join(ROOT, "imagequilting.jl_amd", "julia", "IQHip.jl")).read()
    call = src[src.index("desc = PlanDesc("):]
    call = call[:call.index("pointer(disabled)")]
    fns = re.findall(r"(pad3\w*)\(", call)
    defs = dict(re.findall(r"^(pad3\w*)\(t::NTuple\{N,Int\}\) where \{N\} = ntuple\(i -> i <= N \? Int64\(t\[i\]\) : "
                           r"Int64\((\d)\), 3\)", src, re.M))
    return [int(defs[f]) for f in fns]  # fill value of each padded field, in PlanDesc order


def _desc_like_binding(fills, T, tile, sim, ov):
    import ctypes
    geo = prep.geometry(T.shape, tile, sim, ov)
    fields = [T.shape, tile, geo["ovlsize"], geo["spacing"], geo["ntiles"], geo["padsize"]]
    pad = lambda t, f: list(t) + [f] * (3 - len(t))
    d = iqhip._lib.PlanDesc()
    d.ti = T.ctypes.data
    d.ti_dtype = iqhip._lib.IQ_F64
    for name, vals, f in zip(("tisize", "tilesize", "ovlsize", "spacing", "ntiles", "padsize"), fields, fills):
        getattr(d, name)[:] = pad(vals, f)
    nt = int(np.prod(geo["ntiles"]))
    keep = [np.zeros(int(np.prod(geo["distsize"])), np.uint8), np.arange(nt, dtype=np.int64), np.zeros(nt, np.uint8)]
    d.disabled, d.path, d.skipped = (a.ctypes.data for a in keep)
    d.npath, d.naux, d.tol, d.nreal = nt, 0, 0.1, 1
    return d, keep


def test_julia_binding_descriptor_passes_plan_checks():
    # a 2-D call through IQHip.jl: trailing axis one voxel thick, no overlap
    import ctypes
    L = iqhip._lib.load()
    fills = _binding_pad_functions()
    assert fills == [1, 1, 0, 1, 1, 1]  # tisize, tilesize, ovlsize, spacing, ntiles, padsize
    T = np.zeros((40, 30))
    for tile, sim, ov in (((10, 10), (50, 45), (1 / 6, 1 / 6)), ((30, 30), (40, 30), (0.3, 0.5))):
        d, keep = _desc_like_binding(fills, T, tile, sim, ov)
        assert L.iq_plan_validate(ctypes.byref(d), 1) == 0, L.iq_last_error(None)
        # padding the overlap with 1 (one voxel of overlap on a one-voxel
        # axis) breaks spacing = tilesize - ovlsize there
        bad, keep2 = _desc_like_binding([1] * 6, T, tile, sim, ov)
        assert L.iq_plan_validate(ctypes.byref(bad), 1) == -1
        assert b"spacing" in L.iq_last_error(None)


# --------------------------------------- the MFMA scan's interval bound
def _bf16_rne(x):
    """Round FP32 values to bfloat16 (round to nearest even), as FP32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return r.astype(np.uint32).view(np.float32)


def _scan_ab32(a, b, split, order):
    """The scan's FP32 correlation, emulated: split parts (hi, lo) of each
    FP32 value, products hi*hi + hi*lo + lo*hi (exact in FP32), summed in
    FP32 one term at a time in the given order."""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    ah, bh = _bf16_rne(a), _bf16_rne(b)
    if split:
        al, bl = _bf16_rne(a - ah), _bf16_rne(b - bh)
        terms = np.concatenate([ah.astype(np.float64) * bh, ah.astype(np.float64) * bl, al.astype(np.float64) * bh])
    else:
        terms = ah.astype(np.float64) * bh
    assert np.array_equal(terms.astype(np.float32).astype(np.float64), terms)  # products exact in FP32
    terms = terms.astype(np.float32)
    if order == "ascending":
        terms = terms[np.argsort(np.abs(terms))]
    elif order == "descending":
        terms = terms[np.argsort(-np.abs(terms))]
    acc = np.float32(0.0)
    for t in terms:
        acc = np.float32(acc + t)
    return float(acc)


def _adversarial_f32(n, rng, sign=1.0):
    """FP32 values whose bf16 split drops the most: t - hi just below half a
    bf16 ulp (hi rounds down) and t - hi - lo just below half an ulp of lo
    (lo rounds down), so every dropped part has the sign of t."""
    mant = (rng.integers(0, 2, n).astype(np.uint32) << 16) | np.uint32(0x7FBF)
    expo = rng.integers(120, 134, n).astype(np.uint32) << 23
    return (sign * (expo | mant).view(np.float32)).astype(np.float32)


def test_scan_bound_covers_the_bf16_split():
    from fractions import Fraction
    import ctypes
    L = iqhip._lib.load()
    rng = np.random.default_rng(60)

    def bound(nstep, split):
        g = ctypes.c_double(0.0)
        assert L.iq_scan_error_bound(nstep, int(split), ctypes.byref(g)) == 0
        return g.value

    worst = {}
    for nstep in (1, 2, 3, 4):
        V = 16 * nstep
        for kind in ("adversarial", "random"):
            for trial in range(6):
                if kind == "adversarial":
                    a = _adversarial_f32(V, rng)
                    b = a.copy() if trial % 2 == 0 else _adversarial_f32(V, rng)
                else:
                    a = np.abs(rng.standard_normal(V)).astype(np.float32)
                    b = np.abs(rng.standard_normal(V)).astype(np.float32)
                exact = sum(Fraction(float(x)) * Fraction(float(y)) for x, y in zip(a, b))
                mag = float(sum(abs(Fraction(float(x)) * Fraction(float(y))) for x, y in zip(a, b)))
                for split in (True, False):
                    if not split:
                        a, b = _bf16_rne(a), _bf16_rne(b)  # bf16-exact data take one product
                        exact = sum(Fraction(float(x)) * Fraction(float(y)) for x, y in zip(a, b))
                        mag = float(abs(exact))
                    for order in ("sequential", "ascending", "descending"):
                        err = abs(float(Fraction(_scan_ab32(a, b, split, order)) - exact)) / mag
                        assert err <= bound(nstep, split), (nstep, kind, split, order, err)
                        worst[(nstep, split)] = max(worst.get((nstep, split), 0.0), err)
    # the split's dropped terms reach about 2 * 2^-16 of sum|a||b|: a budget
    # of 4 * 2^-18 for them would not cover a one-k-step mask
    old_budget = bound(1, True) - 2.0 ** -15 * (1 + 2.0 ** -19) + 4 * 2.0 ** -18
    assert worst[(1, True)] > old_budget
    assert worst[(1, True)] > 1.9 * 2.0 ** -16
