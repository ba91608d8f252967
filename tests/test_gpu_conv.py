"""Convolutional encoder of the learned proposal (no counterpart in the
reference): the GPU forward and backward kernels against the FP64 CPU oracle
(oracle/learnref.py, itself pinned by central differences in
tests/test_oracle.py), the banks built from the trained network against the
oracle's forward pass, and what training gains: shortlist recall of the
trained network against the untrained one at equal K."""
import numpy as np
import pytest

import iqhip
import learnref
from iqhip import synth

pytestmark = pytest.mark.gpu

TILE = (8, 8, 4)


def _small_plan():
    rng = np.random.default_rng(11)
    T = synth.grf((20, 18, 10), (6, 6, 3), rng, np.float32)
    plan, info = iqhip.make_plan(T, TILE, (20, 20, 10), overlap=(0.25, 0.25, 0.25), tol=0.1, nreal=1,
                                 rng=np.random.default_rng(1))
    return T, plan


def _corners(c, dist):
    c = np.asarray(c)
    return np.stack([c % dist[0], (c // dist[0]) % dist[1], c // (dist[0] * dist[1])], axis=-1)


def _toff(tv, tile):
    tv = np.asarray(tv)
    return np.stack([tv % tile[0], (tv // tile[0]) % tile[1], tv // (tile[0] * tile[1])], axis=-1)


def _groups(V):
    C = learnref.C
    sizes = [("w1", C * 27), ("b1", C), ("w2", C * C * 27), ("b2", C), ("w3", C * 27), ("b3", 1), ("W", 64 * V)]
    o = 0
    for name, n in sizes:
        yield name, slice(o, o + n)
        o += n


def test_conv_gradient_matches_oracle():
    T, plan = _small_plan()
    rng = np.random.default_rng(12)
    dist = [t - s + 1 for t, s in zip(T.shape, TILE)]
    P = int(np.prod(dist))
    tv = plan.mask_voxels(plan.nmasks() - 1)
    mask = plan.nmasks() - 1
    toff = _toff(tv, TILE)
    V = len(toff)
    C = learnref.C
    w1 = rng.normal(0, 0.3, (C, 1, 27))
    b1 = rng.normal(0, 0.1, C)
    w2 = rng.normal(0, 0.06, (C, C, 27))
    b2 = rng.normal(0, 0.05, C)
    w3 = rng.normal(0, 0.05, (1, C, 27))
    b3 = np.array([0.1])
    W = rng.normal(0, 1 / np.sqrt(V), (64, V))
    theta = learnref.pack(w1, b1, w2, b2, w3, b3, W).astype(np.float32)
    nq, ns = 3, 40
    cands = np.stack([rng.choice(P, ns, replace=False) for _ in range(nq)]).astype(np.int32)
    Q = rng.normal(0, 1, (nq, V)).astype(np.float32)
    D = rng.random((nq, ns)) * 8.0
    tau = np.array([1.0, 2.5, 0.6])
    print("conv gradient: TI", T.shape, "mask", mask, "V", V, flush=True)
    loss, grad = plan.conv_grad(mask, theta, Q, cands, D, tau)
    plan.close()
    x = T.astype(np.float64)
    corners = _corners(cands, dist)
    A2 = np.stack([(learnref.windows(x, corners[i], toff) ** 2).sum(axis=1) for i in range(nq)])
    rl, rg, _ = learnref.loss_and_grad(x, theta.astype(np.float64), toff, A2, Q.astype(np.float64), corners, D, tau)
    assert abs(loss - rl) <= 1e-4 * abs(rl) + 1e-6, (loss, rl)
    gmax = np.abs(rg).max()
    for name, sl in _groups(V):
        ref, got = rg[sl], grad[sl].astype(np.float64)
        if name == "b3":
            # a shift of f moves every window encoding alike: the softmax over a
            # query's candidates does not change, so dL/db3 = 0
            assert abs(ref[0]) <= 1e-12 * gmax and abs(got[0]) <= 1e-5 * gmax, (ref, got)
            continue
        scale = np.abs(ref).max()
        assert scale > 0, name
        err = np.abs(got - ref).max() / scale
        print(name, "max |grad|", scale, "rel err", err, flush=True)
        assert err < 2e-3, (name, err)


def _grf_plan(nreal=1, seed=0):
    T = synth.grf((64, 64, 16), (20, 20, 4), np.random.default_rng(40), np.float32)
    plan, info = iqhip.make_plan(T, (12, 12, 4), (72, 72, 16), overlap=(0.25, 0.25, 0.25), tol=0.1, nreal=nreal,
                                 rng=np.random.default_rng(seed))
    return T, plan


def _recall(plan, U, K):
    plan.set_proposal(True, K=K, check_every=1, min_recall=0.0)
    plan.run(U)
    st = plan.proposal_stats()
    plan.set_proposal(False)
    return st["recall"]


def test_trained_network_gains_over_untrained_and_banks_match_oracle():
    T, plan = _grf_plan(nreal=2)
    U = np.random.default_rng(3).random((2, plan.npath))
    # untrained: random orthonormal heads, network at its initialisation (zero output layer)
    plan.learn(nsamples=2048, iters=0, seed=5)
    st0 = plan.learn_conv(nqueries=64, topn=64, nrandom=16, nhard=16, steps=0, seed=7, val_k=48)
    assert st0["kept_step"] == 0
    K = 48
    r_untrained = _recall(plan, U, K)
    plan.learn(nsamples=2048, iters=0, seed=5)
    print("train", flush=True)
    st = plan.learn_conv(nqueries=256, topn=64, nrandom=16, nhard=16, steps=120, batch=8, lr=1e-3, lr_head=3e-3,
                         seed=7, val_k=48)
    print("stats", st, flush=True)
    assert st["pairs"] == st["queries"] * 96 and st["queries"] == 256
    assert st["loss_last"] < st["loss_first"]
    # the kept parameters recall at least as much of the held-out bands as the start
    assert st["heldout_recall_kept"] >= st["heldout_recall_start"] and st["kept_step"] > 0
    r_trained = _recall(plan, U, K)
    print("recall untrained", r_untrained, "trained", r_trained, flush=True)
    assert r_trained > r_untrained + 0.05
    # banks: the network's output map through the trained heads, rounded to bf16
    theta = plan.conv_params()
    x = T.astype(np.float64)
    dist = [t - s + 1 for t, s in zip(T.shape, (12, 12, 4))]
    P = int(np.prod(dist))
    rng = np.random.default_rng(9)
    for m in range(plan.nmasks()):
        W, tv = plan.encoder(m)
        toff = _toff(tv, (12, 12, 4))
        full = np.concatenate([theta.astype(np.float64), W.astype(np.float64).ravel()])
        _, _, f = learnref.forward(x, full, len(toff))
        c = rng.choice(P, 64, replace=False).astype(np.int32)
        ref = learnref.windows(f, _corners(c, dist), toff) @ W.astype(np.float64).T
        got = plan.bank_rows(m, c).astype(np.float64)
        tol = 2.0 ** -7 * np.abs(ref).max(axis=1, keepdims=True) + 1e-6
        assert (np.abs(got - ref) <= tol).all(), m
    plan.close()
