"""Runs replayed from their captured launch sequence (HIP graph) against runs
launched kernel by kernel from the host: every replay, with other runs of
another plan in between, gives the host-launched results bit for bit."""
import os

import numpy as np
import pytest

import iqhip
from iqhip import synth

pytestmark = pytest.mark.gpu


def _plan(graph: bool, T, tile, sim, ov, nreal):
    os.environ["IQHIP_GRAPH"] = "1" if graph else "0"
    try:
        plan, info = iqhip.make_plan(T, tile, sim, overlap=ov, tol=0.1, nreal=nreal, rng=np.random.default_rng(1))
    finally:
        os.environ.pop("IQHIP_GRAPH", None)
    return plan


def _res(plan):
    r = plan.results(grids=True, chosen=True, cuts=True)
    return r["grids"], r["chosen"], r["cuts"]


def _same(a, b):
    ga, ca, ka = a
    gb, cb, kb = b
    assert np.array_equal(ca, cb)
    for x, y in zip(ga, gb):
        assert np.array_equal(x, y)
    for x, y in zip(ka, kb):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("case", ["grf3d", "binary2d"])
def test_graph_replays_equal_host_launches(case):
    if case == "grf3d":
        T = synth.grf((48, 40, 12), (8, 8, 3), np.random.default_rng(21), np.float32)
        args = (T, (12, 12, 4), (60, 60, 12), (0.25, 0.25, 0.25), 2)
    else:
        x = synth.grf((60, 60), (8, 8), np.random.default_rng(22), np.float64)
        T = (x > np.quantile(x, 0.7)).astype(np.float64)
        args = (T, (12, 12), (80, 80), (1 / 6, 1 / 6), 1)
    print("graph replay", case, flush=True)
    eager, graph = _plan(False, *args), _plan(True, *args)
    nreal = args[-1]
    Us = [np.random.default_rng(k).random((nreal, eager.npath)) for k in range(3)]
    ref = []
    for U in Us:
        eager.run(U)
        ref.append(_res(eager))
    for k in [0, 1, 2, 0, 1]:  # capture, then replays
        graph.run(Us[k])
        _same(_res(graph), ref[k])
        eager.run(Us[(k + 1) % 3])  # other work on the device between replays
        _same(_res(eager), ref[(k + 1) % 3])
    eager.close()
    graph.close()


def test_graph_replays_learned_shortlist_without_checks():
    T = synth.grf((64, 64, 16), (20, 20, 4), np.random.default_rng(40), np.float32)
    args = (T, (12, 12, 4), (72, 72, 16), (0.25, 0.25, 0.25), 2)
    eager, graph = _plan(False, *args), _plan(True, *args)
    Us = [np.random.default_rng(10 + k).random((2, eager.npath)) for k in range(2)]
    for p in (eager, graph):
        p.learn(nsamples=2048, iters=6, seed=1)
        p.set_proposal(True, K=256, check_every=0)
    ref = []
    for U in Us:
        eager.run(U)
        ref.append(_res(eager))
    for k in [0, 1, 0]:
        graph.run(Us[k])
        _same(_res(graph), ref[k])
    eager.close()
    graph.close()
