"""Error paths of the C ABI on the device: inputs the reference cannot
simulate return a status (no hang, no fault).  Every kernel here runs to its
normal end; the conditions are made on the host."""
import ctypes

import numpy as np
import pytest

import iqhip
from iqhip import _lib

pytestmark = pytest.mark.gpu

IQ_ERR_NOCAND = -3


def test_all_masked_training_image_is_refused_at_plan_creation():
    # every training-image window holds a NaN: no candidate pattern
    T = np.full((20, 20), np.nan)
    print("all-NaN training image", flush=True)
    with pytest.raises(_lib.IQError) as e:
        iqhip.make_plan(T, (5, 5), (30, 30), overlap=(0.2, 0.2))
    assert e.value.code == IQ_ERR_NOCAND


def test_relaxation_without_finite_distance_returns_nocand(ctx):
    print("relaxation of all-Inf maps", flush=True)
    with pytest.raises(_lib.IQError) as e:
        iqhip.relaxation(np.full(5000, np.inf), [np.full(5000, np.inf)], 0.1, ctx=ctx)
    assert e.value.code == IQ_ERR_NOCAND


def test_band_pick_of_all_infinite_distances_selects_every_index(ctx):
    # iqsim.jl:237: minimum(D) = Inf and every D .<= Inf
    print("band of all-Inf map", flush=True)
    got = iqhip.band_pick(np.full(3000, np.inf), 0.1, ctx=ctx)
    assert np.array_equal(got, np.arange(3000))


def test_learned_run_survives_a_one_candidate_training_image(ctx):
    # a training image with a single enabled window: the learned shortlist
    # holds it whatever the bracket says, and the run equals the full scan
    rng = np.random.default_rng(11)
    T = rng.standard_normal((24, 24, 8)).astype(np.float32)
    mask = np.ones(T.shape, dtype=bool)
    mask[:12, :12, :4] = False  # only the corner window (12 x 12 x 4 at 0,0,0) is fully active
    T[mask] = np.nan
    plan, info = iqhip.make_plan(T, (12, 12, 4), (30, 30, 8), overlap=(0.25, 0.25, 0.25), rng=np.random.default_rng(1),
                                 ctx=ctx)
    U = np.random.default_rng(2).random((1, plan.npath))
    plan.run(U)
    ref = plan.results()
    print("learned proposal, one candidate", flush=True)
    plan.learn(nsamples=64, iters=0, seed=1)
    plan.set_proposal(True, K=1, check_every=0)
    plan.run(U)
    got = plan.results()
    assert np.array_equal(got["chosen"], ref["chosen"]) and np.all(ref["chosen"] == 0)
    assert np.array_equal(got["grids"][0], ref["grids"][0])
    plan.close()
