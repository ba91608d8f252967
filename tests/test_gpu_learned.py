"""Learned proposal (no counterpart in the reference): encoders learned on
the GPU, their subspace against a CPU PCA of the same windows, exactness of
the shortlist plumbing, recall checks, fallback, and what the learned
encoder gains over an untrained one."""
import numpy as np
import pytest

import iqhip
from iqhip import synth

pytestmark = pytest.mark.gpu


def _plan(nreal=1, seed=0, tile=(12, 12, 4), dtype=np.float32):
    T = synth.grf((64, 64, 16), (20, 20, 4), np.random.default_rng(40), np.float32).astype(dtype)
    plan, info = iqhip.make_plan(T, tile, (72, 72, 16), overlap=(0.25, 0.25, 0.25), tol=0.1, nreal=nreal,
                                 rng=np.random.default_rng(seed))
    return T, plan, info


def _run(plan, U):
    plan.run(U)
    r = plan.results(grids=True, chosen=True)
    return r["grids"], r["chosen"]


def test_encoder_matches_cpu_pca():
    T, plan, info = _plan()
    print("learn", flush=True)
    plan.learn(nsamples=4096, iters=12, seed=1)
    rng = np.random.default_rng(2)
    dist = [t - s + 1 for t, s in zip(T.shape, (12, 12, 4))]
    for m in range(plan.nmasks()):
        W, tv = plan.encoder(m)
        assert np.abs(W.astype(np.float64) @ W.T.astype(np.float64) - np.eye(64)).max() < 1e-3
        # fresh windows gathered on the CPU at the mask's voxels
        x, y, z = tv % 12, (tv // 12) % 12, tv // 144
        cs = [rng.integers(0, d, 3000) for d in dist]
        X = np.stack([T[cx + x, cy + y, cz + z] for cx, cy, cz in zip(*cs)]).astype(np.float64)
        _, S, Vt = np.linalg.svd(X, full_matrices=False)
        best = (S[:64] ** 2).sum()
        got = ((X @ W.T.astype(np.float64)) ** 2).sum()
        assert got >= 0.97 * best, (m, got / best)
    plan.close()


@pytest.mark.parametrize("tile,dtype", [((12, 12, 4), np.float32), ((16, 16, 4), np.float32), ((12, 12, 4), np.float64)],
                         ids=["rows3", "rows4", "float64"])
def test_full_shortlist_is_the_full_scan(tile, dtype):
    # K = every candidate: the shortlist path must reproduce the full scan
    # exactly, with overlap masks of rows of 3 and 12 voxels (12-wide tiles)
    # and of 4 and 16 voxels (16-wide tiles), and on a Float64 training image
    # (values exact in FP32), whose FP32 bounds convert both operands
    T, plan, info = _plan(nreal=2, tile=tile, dtype=dtype)
    U = np.random.default_rng(3).random((2, plan.npath))
    g0, c0 = _run(plan, U)
    plan.learn(nsamples=2048, iters=6, seed=1)
    plan.set_proposal(True, K=10 ** 9, check_every=0)
    print("full shortlist", flush=True)
    g1, c1 = _run(plan, U)
    assert np.array_equal(c0, c1)
    for a, b in zip(g0, g1):
        assert np.array_equal(a, b)
    plan.close()


def test_every_batch_checked_is_exact():
    # check_every=1, min_recall=1: the full scan decides every batch
    T, plan, info = _plan()
    U = np.random.default_rng(4).random((1, plan.npath))
    g0, c0 = _run(plan, U)
    plan.learn(nsamples=2048, iters=6, seed=1)
    plan.set_proposal(True, K=16, check_every=1, min_recall=1.0)
    print("checked", flush=True)
    g1, c1 = _run(plan, U)
    st = plan.proposal_stats()
    assert np.array_equal(c0, c1) and np.array_equal(g0[0], g1[0])
    assert st["queries"] > 0 and 0 <= st["in_shortlist"] <= st["band_members"]
    plan.close()


def test_sampled_checks():
    # check batches that hand only their first queries to the full scan:
    # recall is measured on those alone, and with the whole pattern database
    # as the shortlist the run still equals the full scan bit for bit
    T, plan, info = _plan(nreal=2)
    U = np.random.default_rng(8).random((2, plan.npath))
    g0, c0 = _run(plan, U)
    plan.learn(nsamples=2048, iters=6, seed=1)
    plan.set_proposal(True, K=64, check_every=1, min_recall=0.0, check_queries=0)
    plan.run(U)
    every = plan.proposal_stats()
    plan.set_proposal(True, K=64, check_every=1, min_recall=0.0, check_queries=3)
    print("sampled checks", flush=True)
    plan.run(U)
    sampled = plan.proposal_stats()
    print("checked queries: all", every["queries"], "sampled", sampled["queries"], flush=True)
    assert 0 < sampled["queries"] < every["queries"]
    assert 0 <= sampled["in_shortlist"] <= sampled["band_members"]
    plan.set_proposal(True, K=10 ** 9, check_every=2, min_recall=1.0, check_queries=3)
    g1, c1 = _run(plan, U)
    st = plan.proposal_stats()
    assert not st["fell_back"] and st["in_shortlist"] == st["band_members"] > 0
    assert np.array_equal(c0, c1)
    for a, b in zip(g0, g1):
        assert np.array_equal(a, b)
    plan.close()


def test_fallback_when_recall_drops():
    T, plan, info = _plan()
    U = np.random.default_rng(5).random((1, plan.npath))
    g0, c0 = _run(plan, U)
    plan.learn(nsamples=2048, iters=0, seed=1)  # untrained encoder, tiny shortlist
    plan.set_proposal(True, K=1, check_every=4, min_recall=1.0)
    print("fallback", flush=True)
    g1, c1 = _run(plan, U)
    st = plan.proposal_stats()
    print("fallback stats", st, flush=True)
    # a one-member shortlist cannot hold a band of several members: the
    # checked batches see recall below 1, and the run returns to the full scan
    assert st["queries"] > 0 and st["band_members"] > st["queries"]
    assert st["in_shortlist"] < st["band_members"]
    assert st["fell_back"]
    plan.close()


def test_learned_encoder_gains_over_untrained():
    # same shortlist size, every batch checked: recall of the learned
    # encoder against a random orthonormal one (the method without learning)
    rec = {}
    for iters in (0, 12):
        T, plan, info = _plan(seed=6)
        plan.learn(nsamples=4096, iters=iters, seed=1)
        plan.set_proposal(True, K=64, check_every=1, min_recall=0.0)
        U = np.random.default_rng(7).random((1, plan.npath))
        print("recall iters", iters, flush=True)
        plan.run(U)
        rec[iters] = plan.proposal_stats()["recall"]
        plan.close()
    print("recall untrained %.4f learned %.4f" % (rec[0], rec[12]), flush=True)
    assert rec[12] > rec[0] and rec[12] >= 0.9
