"""Per-function parity of the HIP entry points against the FP64 oracle.

Every comparison is bit-exact: the oracle and the kernels evaluate the same
FP64 operation sequence (fused-multiply-add chains in ascending column-major
order for the distance maps, the reference's operator order elsewhere).
"""
import numpy as np
import pytest

import iqhip
import iqref

pytestmark = pytest.mark.gpu


def _smooth(rng, shape, width=5, dtype=np.float32):
    x = rng.standard_normal(shape)
    k = np.ones(width) / width
    for ax in range(len(shape)):
        if shape[ax] > width:
            x = np.apply_along_axis(lambda v: np.convolve(v, k, mode="same"), ax, x)
    return x.astype(dtype)


# ---------------------------------------------------------------- imfilter
@pytest.mark.parametrize("ishape,kshape", [((200, 100), (30, 10)), ((50, 100, 150), (10, 20, 30)), ((7, 5, 3), (7, 1, 2))])
def test_imfilter_matches_oracle(ctx, ishape, kshape):
    # sizes of runtests.jl:146-161 plus a degenerate one
    print("imfilter", ishape, kshape, flush=True)
    rng = np.random.default_rng(11)
    img, krn = rng.random(ishape), rng.random(kshape)
    got = iqhip.imfilter(img, krn, ctx=ctx)
    ref = iqref.imfilter(img, krn)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


# ------------------------------------------------------------ fastdistance
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("ishape,kshape,density", [((40, 30), (9, 7), 0.3), ((24, 20, 12), (6, 5, 4), 0.5),
                                                  ((24, 20, 12), (6, 5, 4), 0.0), ((30, 30), (8, 8), 1.0)])
def test_fastdistance_matches_oracle(ctx, dtype, ishape, kshape, density):
    print("fastdistance", dtype.__name__, ishape, kshape, density, flush=True)
    rng = np.random.default_rng(5)
    img = _smooth(rng, ishape, dtype=dtype)
    kern = rng.standard_normal(kshape).astype(dtype).astype(np.float64)
    w = rng.random(kshape) < density
    got = iqhip.fastdistance(img, kern, w, ctx=ctx)
    ref = iqref.fastdistance(img.astype(np.float64), kern, w)
    assert np.array_equal(got, ref)


# --------------------------------------------------------------- band pick
def test_band_pick_matches_oracle(ctx):
    rng = np.random.default_rng(3)
    for trial in range(20):
        n = int(rng.integers(1, 5000))
        D = np.round(rng.random(n) * 50) / 7.0
        D[rng.random(n) < 0.2] = np.inf
        if trial == 0:
            D[:] = 0.0
        tol = float(rng.choice([0.1, 0.05, 1.0, 0.3]))
        got = iqhip.band_pick(D, tol, ctx=ctx)
        ref = iqref.band(D, tol)
        assert np.array_equal(got, ref), trial


# -------------------------------------------------------------- relaxation
def test_relaxation_matches_oracle(ctx):
    rng = np.random.default_rng(4)
    for trial in range(12):
        n = int(rng.integers(50, 4000))
        naux = int(rng.integers(1, 4))
        D = np.round(rng.random(n) * 30)
        Ds = [np.round(rng.random(n) * 20) for _ in range(naux)]
        dis = rng.random(n) < 0.1
        D[dis] = np.inf
        for d in Ds:
            d[dis] = np.inf
        if trial == 1:
            D[~dis] = 0.0
        tol = float(rng.choice([0.1, 0.05, 0.5]))
        got = iqhip.relaxation(D, Ds, tol, ctx=ctx)
        ref = iqref.relaxation(D, Ds, tol)
        assert np.array_equal(got, ref), trial


@pytest.mark.parametrize("kind", ["continuous", "ties", "allzero", "mixed_inf"])
def test_relaxation_large_matches_oracle(ctx, kind):
    """Pattern databases of 300k candidates: each prefix set comes from an
    order statistic over the whole map (histogram passes, then a bin ranked
    in LDS, or the tied value's members in index order)."""
    rng = np.random.default_rng({"continuous": 1, "ties": 2, "allzero": 3, "mixed_inf": 4}[kind])
    n = 300_001  # not a multiple of 64
    if kind == "continuous":
        D = rng.random(n) * 1e3
        Ds = [rng.random(n) * 50.0, rng.gamma(2.0, 3.0, n)]
    elif kind == "ties":
        D = np.round(rng.random(n) * 2)
        Ds = [np.round(rng.random(n) * 3), np.zeros(n)]
    elif kind == "allzero":
        D = np.zeros(n)
        Ds = [rng.random(n)]
    else:
        D = rng.random(n) * 10
        Ds = [rng.random(n) * 10]
        dis = rng.random(n) < 0.3
        D[dis] = np.inf
        Ds[0][dis | (rng.random(n) < 0.05)] = np.inf
        D[rng.random(n) < 0.2] = 0.0
    for tol in (0.1, 0.01):
        got = iqhip.relaxation(D, Ds, tol, ctx=ctx)
        ref = iqref.relaxation(D, Ds, tol)
        assert np.array_equal(got, ref), (kind, tol, len(got), len(ref))


# ---------------------------------------------------------------- taumodel
def test_taumodel_matches_oracle_bitwise(ctx):
    rng = np.random.default_rng(8)
    for trial in range(12):
        n = int(rng.integers(2, 3000))
        naux = int(rng.integers(0, 3))
        D = np.round(rng.random(n) * 40) / 3
        Ds = [np.round(rng.random(n) * 10) for _ in range(naux)]
        ev = np.sort(rng.choice(n, size=int(rng.integers(1, min(n, 1500))), replace=False))
        got = iqhip.taumodel(ev, D, Ds, ctx=ctx)
        ref = iqref.taumodel(ev, D, Ds)
        assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), trial


def test_taumodel_large_event_sets_bitwise(ctx):
    # more events than the LDS ranker holds: radix-sorted dense ranks
    rng = np.random.default_rng(18)
    n = 60000
    D = rng.random(n) * 5.0                          # distinct values
    T = np.round(rng.random(n) * 30) - 10.0          # ties, negatives
    T[rng.choice(n, 50, replace=False)] = -0.0       # -0 and +0 tie
    T[rng.choice(n, 50, replace=False)] = 0.0
    C = np.full(n, 2.5)                              # constant source: neutral
    for ev in (np.arange(n), np.sort(rng.choice(n, 9000, replace=False))):
        for Ds in ([], [T], [T, C, D[::-1].copy()]):
            got = iqhip.taumodel(ev, D, Ds, ctx=ctx)
            ref = iqref.taumodel(ev, D, Ds)
            assert np.array_equal(got.view(np.uint64), ref.view(np.uint64)), (len(ev), len(Ds))


# ------------------------------------------------------------------ sample
def test_sample_matches_oracle(ctx):
    rng = np.random.default_rng(9)
    for n in (1, 2, 63, 64, 65, 4095, 4096, 4097, 300000):
        p = rng.random(n)
        for u in (0.0, 0.5, 0.999999, float(rng.random())):
            assert iqhip.sample(p, u, ctx=ctx) == iqref.sample(p, u), (n, u)


# ---------------------------------------------------------------- graphcut
def test_graphcut_reference_convention(ctx):
    # runtests.jl:112-119
    A = np.ones((20, 20))
    B = np.ones((20, 20))
    C = iqhip.graphcut(A, B, 0, ctx=ctx)
    assert C[:-1, :].all() and not C[-1, :].any()
    C = iqhip.graphcut(A, B, 1, ctx=ctx)
    assert C[:, :-1].all() and not C[:, -1].any()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape,dim", [((5, 30), 0), ((30, 5), 1), ((4, 20, 5), 0), ((20, 4, 5), 1), ((6, 24, 6), 0),
                                       ((24, 24, 2), 2), ((8, 48), 0), ((3, 7, 4), 0), ((8, 32, 8), 0),
                                       ((2, 9, 9), 0)])
def test_graphcut_matches_oracle(ctx, dtype, shape, dim):
    print("graphcut", dtype.__name__, shape, dim, flush=True)
    rng = np.random.default_rng(21)
    for trial in range(3):
        if trial == 2:  # categorical codes: many tied capacities
            A = rng.integers(0, 3, shape).astype(dtype)
            B = rng.integers(0, 3, shape).astype(dtype)
        else:
            A = _smooth(rng, shape, 3, dtype)
            B = _smooth(rng, shape, 3, dtype)
        got = iqhip.graphcut(A, B, dim, ctx=ctx)
        ref = iqref.graphcut(A, B, dim)
        assert np.array_equal(got, ref), trial
