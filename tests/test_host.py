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


def test_genpath_matches_oracle():
    for kind in ("raster", "random", "dilation"):
        a = prep.genpath(np.random.default_rng(5), (6, 5, 4), kind, [])
        b = iqref.genpath(np.random.default_rng(5), (6, 5, 4), kind, [])
        assert a == b and sorted(a) == list(range(120))
    a = prep.genpath(np.random.default_rng(5), (6, 5), "raster", [3, 17, 29])
    b = iqref.genpath(np.random.default_rng(5), (6, 5), "raster", [3, 17, 29])
    assert a == b and set(a[:3]) == {3, 17, 29}


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
