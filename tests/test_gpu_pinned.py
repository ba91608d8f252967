"""Plan results copied into page-locked host memory (iq_host_alloc, no
counterpart in the reference) equal those copied into a pageable array, on
two runs in a row."""
import numpy as np
import pytest

import iqhip
from iqhip import synth

pytestmark = pytest.mark.gpu


def test_results_into_pinned_buffer():
    T = synth.grf((48, 48, 12), (10, 10, 3), np.random.default_rng(31), np.float32)
    plan, info = iqhip.make_plan(T, (16, 16, 4), (60, 60, 12), overlap=(0.25, 0.25, 0.25), tol=0.1, nreal=2,
                                 rng=np.random.default_rng(1))
    n = sum(int(np.prod(p)) for p in plan.real_padsize)
    buf = plan.ctx.pinned_empty(n, plan.dtype)
    assert buf.dtype == plan.dtype and buf.size == n and buf.flags.c_contiguous
    for seed in (2, 3):
        U = np.random.default_rng(seed).random((2, plan.npath))
        print("pinned run", seed, flush=True)
        plan.run(U)
        ref = plan.results(grids=True, chosen=False)["grids"]
        got = plan.results(grids=True, chosen=False, out=buf)["grids"]
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
    del got, buf
    plan.close()
