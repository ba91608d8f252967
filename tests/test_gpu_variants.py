"""Alternative kernels chosen by environment knobs give the default kernels'
results bit for bit: the exact recheck with a wavefront per superset member
(IQHIP_RECHECK=groups; the default runs a lane per member), the min cut of
slabs with more than 1,024 interior nodes on 512 work-items
(IQHIP_CUT_TEAM=512; the default uses 1,024) and the MFMA scan with 3 or 2
wavefronts per group (IQHIP_SCAN_NW; the default is timed per plan).  Tile 32x32x8 with overlap 1/4
gives 6 x 32 x 8 = 1,536-node slabs, the largest the large-slab kernel
takes."""
import os

import numpy as np
import pytest

import iqhip
from iqhip import synth

pytestmark = pytest.mark.gpu


def _run(env, T, U):
    saved = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        plan, info = iqhip.make_plan(T, (32, 32, 8), (56, 56, 14), overlap=(0.25, 0.25, 0.25), tol=0.1, nreal=2,
                                     rng=np.random.default_rng(1))
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    plan.run(U)
    r = plan.results(grids=True, chosen=True, cuts=True)
    rounds, levels, cuts = plan.cut_stats()
    plan.close()
    return r["grids"], r["chosen"], r["cuts"], (rounds, cuts)


@pytest.mark.parametrize("env", [{"IQHIP_RECHECK": "groups"}, {"IQHIP_CUT_TEAM": "512"}, {"IQHIP_SCAN_NW": "3"},
                                 {"IQHIP_SCAN_NW": "2"}], ids=["recheck", "cut512", "scan3", "scan2"])
def test_alternative_kernels_match_default(env):
    T = synth.grf((48, 48, 12), (10, 10, 3), np.random.default_rng(31), np.float32)
    probe, _ = iqhip.make_plan(T, (32, 32, 8), (56, 56, 14), overlap=(0.25, 0.25, 0.25), tol=0.1, nreal=2,
                               rng=np.random.default_rng(1))
    U = np.random.default_rng(32).random((2, probe.npath))
    probe.close()
    print("variant", env, flush=True)
    ga, ca, ka, sa = _run({}, T, U)
    gb, cb, kb, sb = _run(env, T, U)
    assert sa[1] > 0 and sa == sb  # same cuts, same push rounds
    assert np.array_equal(ca, cb)
    for x, y in zip(ga, gb):
        assert np.array_equal(x, y)
    for x, y in zip(ka, kb):
        assert np.array_equal(x, y)
