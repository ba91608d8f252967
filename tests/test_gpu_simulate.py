"""End-to-end parity of the device tile loop (iq_plan_run) with the oracle.

Both sides receive the same uniforms, so every pattern pick, every cut and
every pasted voxel must agree exactly.
"""
import math

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


def _compare(T, tile, sim, seed=0, **kw):
    (reals, cuts, voxs), tr = iqhip.iqsim(T, tile, sim, rng=np.random.default_rng(seed), debug=True,
                                          return_trace=True, **kw)
    # the oracle replays the product's path; genpath itself is checked in test_host.py
    (rr, rc, rv), rtr = iqref.iqsim(T, tile, sim, uniforms=tr["uniforms"], debug=True, return_trace=True,
                                    simpath=tr["path"], **kw)
    assert list(tr["path"]) == list(rtr["path"])
    assert np.array_equal(tr["chosen"], rtr["chosen"])
    for a, b in zip(reals, rr):
        a = np.asarray(a, dtype=object if a.dtype == object else a.dtype)
        if a.dtype == object:
            assert all((x is None and y is None) or x == y for x, y in zip(a.ravel(), np.asarray(b).ravel()))
        else:
            assert np.array_equal(a, b, equal_nan=True)
    for a, b in zip(cuts, rc):
        assert np.array_equal(a, b, equal_nan=True)
    assert voxs == rv
    tr["nband"] = rtr["nband"]
    return reals, tr


def test_sim_2d_binary_f64():
    rng = np.random.default_rng(1)
    x = _smooth(rng, (60, 60), 7, np.float64)
    T = (x > np.quantile(x, 0.7)).astype(np.float64)
    print("2d binary", flush=True)
    _compare(T, (15, 15), (80, 80), tol=0.1)


def test_sim_3d_continuous_f32():
    rng = np.random.default_rng(2)
    T = _smooth(rng, (30, 30, 12), 5, np.float32)
    print("3d continuous f32 (z overlap of one voxel: no z cut)", flush=True)
    with pytest.warns(UserWarning):
        _compare(T, (10, 10, 5), (40, 40, 12), overlap=(1 / 5, 1 / 5, 1 / 5), tol=0.1)


def test_sim_3d_full_overlap_f64_nreal():
    rng = np.random.default_rng(3)
    T = _smooth(rng, (20, 20, 16), 3, np.float64)
    print("3d f64 nreal=2", flush=True)
    _compare(T, (8, 8, 8), (20, 20, 20), overlap=(1 / 4, 1 / 4, 1 / 4), nreal=2)


def test_sim_categorical_int():
    rng = np.random.default_rng(4)
    x = _smooth(rng, (30, 30, 8), 5, np.float64)
    T = np.digitize(x, np.quantile(x, [0.65, 0.9])).astype(np.int64)
    print("categorical", flush=True)
    reals, _ = _compare(T, (10, 10, 4), (35, 35, 8), overlap=(1 / 4, 1 / 4, 1 / 4))
    assert set(v for v in reals[0].ravel()) <= set(np.unique(T).tolist())


def test_sim_soft_and_hard():
    rng = np.random.default_rng(5)
    T = _smooth(rng, (40, 40), 5, np.float32)
    aux_ti = _smooth(rng, (40, 40), 3, np.float32)
    aux = _smooth(rng, (50, 50), 3, np.float32)
    hard = {(int(i), int(j)): float(T[i % 40, j % 40]) for i, j in zip(rng.integers(0, 50, 12), rng.integers(0, 50, 12))}
    print("soft + hard", flush=True)
    _compare(T, (10, 10), (50, 50), soft=[(aux, aux_ti)], hard=hard, tol=0.2)


@pytest.mark.parametrize("bx", ["32", "16"])
def test_sim_soft_and_hard_3d_map_blocks(monkeypatch, bx):
    # exact multi-source maps with 32 x 16 and 16 x 16 candidate blocks
    # (IQHIP_MAPS_BX=16 forces the narrow ones, which plans take when the
    # wide block's region does not fit in LDS); candidate grid 27 x 23 x 5
    monkeypatch.setenv("IQHIP_MAPS_BX", bx)
    rng = np.random.default_rng(25)
    T = _smooth(rng, (36, 30, 8), 4, np.float32)
    aux_ti = _smooth(rng, (36, 30, 8), 3, np.float32)
    aux = _smooth(rng, (30, 30, 10), 3, np.float32)
    hard = {(int(i), int(j), int(k)): float(T[i % 36, j % 30, k % 8])
            for i, j, k in zip(rng.integers(0, 30, 10), rng.integers(0, 30, 10), rng.integers(0, 10, 10))}
    print("soft + hard 3-D, map blocks", bx, flush=True)
    _compare(T, (10, 8, 4), (30, 30, 10), overlap=(0.3, 0.25, 0.25), soft=[(aux, aux_ti)], hard=hard, tol=0.2)


def test_sim_masked_ti_and_grid():
    T = np.ones((20, 20, 20))
    T[:, 4, :] = np.nan
    shape = {(i, 4, k): float("nan") for i in range(20) for k in range(20)}
    print("masked ti + grid", flush=True)
    _compare(T, (10, 10, 10), (20, 20, 20), hard=shape)


@pytest.mark.parametrize("path", ["random", "dilation"])
def test_sim_nonraster_paths(path):
    # general wave construction: tiles run as soon as their earlier-path neighbours are pasted
    rng = np.random.default_rng(6)
    T = _smooth(rng, (30, 30, 10), 5, np.float32)
    print("path", path, flush=True)
    _compare(T, (10, 10, 5), (40, 40, 15), overlap=(0.3, 0.3, 0.4), path=path, seed=9)


def test_sim_data_first_path_2d():
    rng = np.random.default_rng(7)
    T = _smooth(rng, (50, 50), 5, np.float64)
    hard = {(3, 4): 0.5, (40, 7): -0.25, (20, 33): 1.0}
    print("data-first", flush=True)
    _compare(T, (12, 12), (55, 55), hard=hard, seed=4)


def test_sim_large_overlap_runs_sequentially():
    # overlap above one half: tiles two apart overlap, every tile is its own wave
    rng = np.random.default_rng(8)
    T = _smooth(rng, (30, 30), 5, np.float32)
    print("overlap 0.6", flush=True)
    _compare(T, (10, 10), (30, 30), overlap=(0.6, 0.6))


def test_sim_valu_scan_path(monkeypatch):
    # the FP32 VALU scan (selected with IQHIP_SCAN=valu) against the oracle
    monkeypatch.setenv("IQHIP_SCAN", "valu")
    rng = np.random.default_rng(10)
    T = _smooth(rng, (30, 30, 12), 5, np.float32)
    print("valu scan", flush=True)
    _compare(T, (10, 10, 4), (40, 40, 12), overlap=(0.25, 0.25, 0.25))


@pytest.mark.parametrize("nw", ["2", "4"])
def test_sim_scan_group_sizes(monkeypatch, nw):
    # the MFMA scan with 2- and with 4-wavefront groups (forced with
    # IQHIP_SCAN_NW; the default is measured per plan) against the oracle
    monkeypatch.setenv("IQHIP_SCAN_NW", nw)
    rng = np.random.default_rng(20)
    T = _smooth(rng, (44, 40, 14), 5, np.float32)
    print("scan groups of", nw, "wavefronts", flush=True)
    _compare(T, (10, 10, 4), (40, 40, 14), overlap=(0.25, 0.25, 0.25))


def test_scan_tuning_reports_a_group_size():
    rng = np.random.default_rng(21)
    T = _smooth(rng, (40, 40, 12), 5, np.float32)
    plan, _ = iqhip.make_plan(T, (10, 10, 4), (60, 60, 16), overlap=(0.25, 0.25, 0.25), nreal=2)
    (n1, n2), us = plan.scan_tuning()
    assert n1 in (2, 4) and n2 in (2, 4)
    assert all(v > 0 for v in us.values())


def test_sim_large_tied_band():
    # binary noise, random path (overlaps on several sides from unrelated
    # windows): bands of thousands of candidates over a few distinct integer
    # distances (radix-ranked beyond the LDS ranker)
    rng = np.random.default_rng(11)
    T = (rng.random((140, 140)) > 0.5).astype(np.float64)
    print("large tied band", flush=True)
    _, tr = _compare(T, (10, 10), (46, 46), overlap=(0.3, 0.3), tol=1.0, path="random")
    nb = tr["nband"]
    assert ((nb > 4096) & (nb < 131 * 131)).any()


def test_sim_large_continuous_band():
    rng = np.random.default_rng(12)
    T = rng.standard_normal((120, 120)).astype(np.float32)
    print("large continuous band", flush=True)
    _, tr = _compare(T, (10, 10), (46, 46), overlap=(0.3, 0.3), tol=1.0, path="random")
    nb = tr["nband"]
    assert ((nb > 4096) & (nb < 111 * 111)).any()


def test_sim_soft_large_database():
    # first tile: D == 0, the database is the first tenth of the soft order
    # (> 4096 events with distinct soft distances)
    rng = np.random.default_rng(13)
    T = _smooth(rng, (230, 230), 5, np.float32)
    aux_ti = _smooth(rng, (230, 230), 3, np.float32)
    aux = _smooth(rng, (40, 40), 3, np.float32)
    print("soft large database", flush=True)
    _, tr = _compare(T, (10, 10), (40, 40), soft=[(aux, aux_ti)], tol=0.1)
    assert (tr["nband"] > 4096).any()


def _compare_batch(T, tile, sizes, nreal, seed=0, **kw):
    outs, trs = iqhip.iqsim_batch(T, tile, sizes, nreal=nreal, rng=np.random.default_rng(seed), debug=True,
                                  return_trace=True, **kw)
    hard = kw.pop("hard", None)
    for g, (size, out, tr) in enumerate(zip(sizes, outs, trs)):
        kg = dict(kw)
        if hard is not None:
            kg["hard"] = hard[g]
        (rr, rc, rv), rtr = iqref.iqsim(T, tile, size, nreal=nreal[g], uniforms=tr["uniforms"], debug=True,
                                        return_trace=True, simpath=tr["path"], **kg)
        assert np.array_equal(tr["chosen"], rtr["chosen"]), g
        reals, cuts, voxs = out
        for a, b in zip(reals, rr):
            assert np.array_equal(np.asarray(a, dtype=object), np.asarray(b, dtype=object)), g
        for a, b in zip(cuts, rc):
            assert np.array_equal(a, b, equal_nan=True), g
        assert voxs == rv
    return outs


def test_sim_batch_uneven_grids_categorical():
    # one plan, four output grids of different sizes, lockstep waves
    rng = np.random.default_rng(14)
    x = _smooth(rng, (30, 30, 8), 5, np.float64)
    T = np.digitize(x, np.quantile(x, [0.65, 0.9])).astype(np.uint8)
    sizes = [(35, 35, 8), (26, 44, 8), (50, 30, 8), (20, 20, 8)]
    print("batch of uneven grids", flush=True)
    outs = _compare_batch(T, (10, 10, 4), sizes, [1, 2, 1, 1], overlap=(1 / 4, 1 / 4, 1 / 4))
    for size, reals in zip(sizes, outs):
        assert reals[0][0].shape == size


def test_sim_batch_soft_and_hard_per_grid():
    rng = np.random.default_rng(15)
    T = _smooth(rng, (40, 40), 5, np.float32)
    aux_ti = _smooth(rng, (40, 40), 3, np.float32)
    aux = _smooth(rng, (60, 60), 3, np.float32)
    sizes = [(50, 50), (36, 58)]
    hard = [{(3, 4): 0.5, (30, 20): -0.3}, {(10, 50): 0.1}]
    print("batch soft + hard", flush=True)
    _compare_batch(T, (10, 10), sizes, [1, 1], soft=[(aux, aux_ti)], hard=hard, tol=0.2)


def test_sim_large_slabs_512_lane_cut():
    # x/y slabs of 8 x 24 x 8 voxels: 1,152 interior cut nodes (512-lane cut)
    rng = np.random.default_rng(16)
    T = _smooth(rng, (44, 44, 12), 5, np.float32)
    print("large slabs", flush=True)
    _compare(T, (24, 24, 8), (40, 40, 10), overlap=(1 / 3, 1 / 3, 1 / 3))
