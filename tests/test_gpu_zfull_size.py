"""Parity at full benchmark sizes: the 3-D Float32 GRF 100x100x50 ->
200x200x50 simulation (tile 20x20x5, overlap 1/5, tol 0.1; 2,197 tile steps
in 85 waves) and the 2-D binary 200x200 -> 400x400 one (tile 30x30, overlap
1/6; 256 steps) against the oracle, bit for bit, plus run-to-run
determinism of the device path.  The oracle's share of the 3-D case takes
about 25 s on 16 host threads.
"""
import numpy as np
import pytest

import iqhip
import iqref
from iqhip import synth

pytestmark = pytest.mark.gpu


def test_full_size_grf3d_matches_oracle():
    w = synth.workload(2)
    T, tile, sim = w["trainimg"], w["tilesize"], w["simsize"]
    kw = dict(overlap=w["overlap"], tol=w["tol"])
    print("full size 100x100x50 -> 200x200x50", flush=True)
    (reals, cuts, voxs), tr = iqhip.iqsim(T, tile, sim, rng=np.random.default_rng(3), debug=True,
                                          return_trace=True, **kw)
    (rr, rc, rv), rtr = iqref.iqsim(T, tile, sim, uniforms=tr["uniforms"], debug=True, return_trace=True,
                                    simpath=tr["path"], **kw)
    assert np.array_equal(tr["chosen"], rtr["chosen"])
    assert np.array_equal(reals[0], rr[0], equal_nan=True)
    assert np.array_equal(cuts[0], rc[0], equal_nan=True)
    assert voxs == rv
    # the same uniforms give the same simulation again (no order-dependent
    # reductions or races on the device path)
    (reals2, cuts2, _), tr2 = iqhip.iqsim(T, tile, sim, rng=np.random.default_rng(3), debug=True,
                                          return_trace=True, **kw)
    assert np.array_equal(tr2["chosen"], tr["chosen"])
    assert np.array_equal(reals2[0], reals[0], equal_nan=True)
    assert np.array_equal(cuts2[0], cuts[0], equal_nan=True)


def test_full_size_binary2d_matches_oracle():
    w = synth.workload(1)
    T, tile, sim = w["trainimg"], w["tilesize"], w["simsize"]
    kw = dict(overlap=w["overlap"], tol=w["tol"])
    print("full size 200x200 binary -> 400x400", flush=True)
    (reals, cuts, voxs), tr = iqhip.iqsim(T, tile, sim, rng=np.random.default_rng(4), debug=True,
                                          return_trace=True, **kw)
    (rr, rc, rv), rtr = iqref.iqsim(T, tile, sim, uniforms=tr["uniforms"], debug=True, return_trace=True,
                                    simpath=tr["path"], **kw)
    assert np.array_equal(tr["chosen"], rtr["chosen"])
    assert np.array_equal(reals[0], rr[0], equal_nan=True)
    assert np.array_equal(cuts[0], rc[0], equal_nan=True)
    assert voxs == rv
