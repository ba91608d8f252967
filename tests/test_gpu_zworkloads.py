"""Parity at the shapes of the batched, soft + hard and large-candidate
workloads, against the oracle, plus property checks of the full runs.

* Batched (3-facies UInt8 channels 150x150x30, 64 grids U{100..300}^2 x 30,
  tile 24x24x6, overlap 1/4): all 64 grids in one lockstep plan (bf16 MFMA
  scan, 403,225 candidates) with properties checked; the smallest grid (whole
  path) and the largest (its first 300 raster steps, two z layers) in one
  batched plan bit-exact against the oracle.
* Soft + hard (2-D Float32 GRF 512^2 -> 1024^2, box-blurred soft pair, 500
  hard points, tile 48^2, overlap 1/6, tol 0.05): the first 250 steps of the
  data-first path of one realization bit-exact against the oracle (hard
  tiles, next slabs, tiled exact maps and the relaxation at 216,225
  candidates); 32 realizations on the device with the hard data honoured.
* Large candidate set (3-D Float32 GRF 256x256x64 -> 384x384x64, tile
  32x32x8, overlap 1/4, 2,885,625 candidates): the first 20 raster steps
  (x and y slabs of 1,536-node cuts) bit-exact against the oracle; the
  learned proposal on all 16 full realizations: recall >= 0.99 without
  falling back, and with every batch checked it reproduces the full scan.

The oracle runs on the host (OpenMP) and takes most of the time here.
"""
import numpy as np
import pytest

import iqhip
import iqref
from iqhip import synth

pytestmark = pytest.mark.gpu


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == object or b.dtype == object:
        return a.shape == b.shape and all((x is None and y is None) or x == y for x, y in zip(a.ravel(), b.ravel()))
    return np.array_equal(a, b, equal_nan=True)


def _oracle(T, tile, sim, U, path, **kw):
    (rr, rc, rv), rtr = iqref.iqsim(T, tile, sim, uniforms=U, simpath=path, debug=True, return_trace=True, **kw)
    return rr, rc, rtr


# ------------------------------------------------------------- batched
def test_batched_channels_smallest_and_largest_grid_match_oracle():
    w = synth.workload(3)
    T, tile, ov, tol = w["trainimg"], w["tilesize"], w["overlap"], w["tol"]
    sizes = w["simsizes"]
    vol = [s[0] * s[1] for s in sizes]
    small, large = sizes[int(np.argmin(vol))], sizes[int(np.argmax(vol))]
    nts = int(np.prod(iqref.geometry(T.shape, tile, small, ov)["ntiles"]))
    paths = [list(range(nts)), list(range(300))]
    rng = np.random.default_rng(31)
    U = [rng.random(len(p)) for p in paths]
    print("batched: grids", small, "(whole path) and", large, "(300 steps)", flush=True)
    outs, traces = iqhip.iqsim_batch(T, tile, [small, large], overlap=ov, tol=tol, nreal=1, debug=True,
                                     uniforms=U, simpaths=paths, return_trace=True)
    for g, sim in enumerate((small, large)):
        (reals, cuts, voxs), tr = outs[g], traces[g]
        print("oracle grid", sim, flush=True)
        rr, rc, rtr = _oracle(T, tile, sim, U[g][None, :], paths[g], overlap=ov, tol=tol)
        assert np.array_equal(tr["chosen"][0], rtr["chosen"][0]), sim
        assert _same(reals[0], rr[0]), sim
        assert np.array_equal(cuts[0], rc[0], equal_nan=True), sim


def test_batched_channels_all_64_grids_properties():
    w = synth.workload(3)
    T, tile, ov, tol = w["trainimg"], w["tilesize"], w["overlap"], w["tol"]
    sizes = w["simsizes"]
    plan, infos = iqhip.make_plan_multi(T, tile, sizes, overlap=ov, tol=tol, nreal=1, rng=np.random.default_rng(2))
    U = [np.random.default_rng(3 + r).random(n) for r, n in enumerate(plan.real_npath)]
    print("batched: 64 grids, lockstep", flush=True)
    plan.run(U)
    r1 = plan.results(grids=True, chosen=True, cuts=True)
    plan.run(U)
    r2 = plan.results(grids=True, chosen=True, cuts=True)
    plan.close()
    codes = set(np.unique(T).tolist())
    assert codes == {0, 1, 2}
    for g, (info, a, b, ca, cb) in enumerate(zip(infos, r1["grids"], r2["grids"], r1["cuts"], r2["cuts"])):
        geo = info["geo"]
        assert a.shape == tuple(geo["padsize"]) and ca.shape == tuple(geo["padsize"])
        assert set(np.unique(a).tolist()) <= codes, g  # facies stay exact codes
        assert set(np.unique(ca).tolist()) <= {0.0, 1.0}, g
        assert np.array_equal(a, b) and np.array_equal(ca, cb), g  # run to run identical
        ch1, ch2 = np.asarray(r1["chosen"][g]), np.asarray(r2["chosen"][g])
        assert np.array_equal(ch1, ch2) and (ch1 >= 0).all()
        # every pick is a window that lies inside the training image
        assert (ch1 < int(np.prod(geo["distsize"]))).all()
        # the realization's output equals the picked windows where no cut kept the grid
        reals, _, _ = iqhip.api.finish(info, [a], [ca])
        assert reals[0].shape == tuple(info["geo"]["simsize"])


# ---------------------------------------------------------- soft + hard
def test_soft_hard_data_first_prefix_matches_oracle():
    w = synth.workload(4)
    T, tile, sim, ov, tol = w["trainimg"], w["tilesize"], w["simsize"], w["overlap"], w["tol"]
    soft, hard = w["soft"], w["hard"]
    geo = iqref.geometry(T.shape, tile, sim, ov)
    skipped, datainds = iqref.findskipped(hard, geo)
    path = iqref.genpath(np.random.default_rng(4), geo["ntiles"], "raster", list(datainds))[:250]
    U = np.random.default_rng(5).random((1, len(path)))
    print("soft + hard: 250 steps of the data-first path,", len(datainds), "data tiles", flush=True)
    (reals, cuts, voxs), tr = iqhip.iqsim(T, tile, sim, overlap=ov, tol=tol, soft=soft, hard=hard, uniforms=U,
                                          simpath=path, debug=True, return_trace=True)
    # next slabs occur on this path prefix
    nt = geo["ntiles"]
    pasted, nexts = set(), 0
    for ind in path:
        t = np.unravel_index(ind, nt, order="F")
        for d in range(2):
            nb = list(t)
            nb[d] += 1
            if nb[d] < nt[d] and np.ravel_multi_index(nb, nt, order="F") in pasted:
                nexts += 1
        pasted.add(ind)
    assert nexts > 0
    print("oracle", flush=True)
    rr, rc, rtr = _oracle(T, tile, sim, U, path, overlap=ov, tol=tol, soft=soft, hard=hard)
    assert np.array_equal(tr["chosen"], rtr["chosen"])
    assert _same(reals[0], rr[0])
    assert np.array_equal(cuts[0], rc[0], equal_nan=True)


def test_soft_hard_32_realizations_honour_hard_data():
    w = synth.workload(4)
    T, tile, sim, ov, tol = w["trainimg"], w["tilesize"], w["simsize"], w["overlap"], w["tol"]
    print("soft + hard: 32 realizations", flush=True)
    reals = iqhip.iqsim(T, tile, sim, overlap=ov, tol=tol, soft=w["soft"], hard=w["hard"], nreal=32,
                        rng=np.random.default_rng(6))
    assert len(reals) == 32
    for r in reals:
        assert r.shape == sim and r.dtype == np.float32 and np.isfinite(r).all()
        for k, v in w["hard"].items():
            assert r[k] == np.float32(v)
    # realizations differ (independent draws)
    assert not np.array_equal(reals[0], reals[1])


# ------------------------------------------------ large candidate set
def test_large_candidate_set_prefix_matches_oracle():
    w = synth.workload(5)
    T, tile, sim, ov, tol = w["trainimg"], w["tilesize"], w["simsize"], w["overlap"], w["tol"]
    path = list(range(20))  # tiles 0-15 of the first row, then 16-19 of the second: x and y slabs
    U = np.random.default_rng(7).random((1, len(path)))
    print("2.9M candidates: first 20 raster steps", flush=True)
    (reals, cuts, voxs), tr = iqhip.iqsim(T, tile, sim, overlap=ov, tol=tol, uniforms=U, simpath=path, debug=True,
                                          return_trace=True)
    print("oracle", flush=True)
    rr, rc, rtr = _oracle(T, tile, sim, U, path, overlap=ov, tol=tol)
    assert np.array_equal(tr["chosen"], rtr["chosen"])
    assert _same(reals[0], rr[0])
    assert np.array_equal(cuts[0], rc[0], equal_nan=True)


def test_learned_proposal_full_size_recall_and_checked_batches():
    w = synth.workload(5)
    T, tile, sim, ov, tol = w["trainimg"], w["tilesize"], w["simsize"], w["overlap"], w["tol"]
    plan, info = iqhip.make_plan(T, tile, sim, overlap=ov, tol=tol, nreal=w["nreal"], rng=np.random.default_rng(8))
    U = np.random.default_rng(9).random((plan.nreal, plan.npath))
    print("full scan, 16 realizations", flush=True)
    plan.run(U)
    full = plan.results(grids=True, chosen=True)
    print("learn", flush=True)
    plan.learn(nsamples=8192, iters=16, seed=1)
    plan.set_proposal(True, K=2048, check_every=16, min_recall=0.99)
    print("learned proposal", flush=True)
    plan.run(U)
    st = plan.proposal_stats()
    print("recall", st, flush=True)
    assert st["queries"] > 0 and st["recall"] >= 0.99 and not st["fell_back"]
    got = plan.results(grids=True, chosen=True)
    assert (np.asarray(got["chosen"]) >= 0).all()
    # every batch checked: the full scan decides, so the run is the full scan's
    plan.set_proposal(True, K=2048, check_every=1, min_recall=0.0)
    print("learned proposal, every batch checked", flush=True)
    plan.run(U)
    chk = plan.results(grids=True, chosen=True)
    plan.close()
    assert np.array_equal(chk["chosen"], full["chosen"])
    for a, b in zip(chk["grids"], full["grids"]):
        assert np.array_equal(a, b)
