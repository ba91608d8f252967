"""iqhip -- MI355X (gfx950) implementation of ImageQuilting.jl's iqsim hot path.

The public functions mirror the reference API (src/ImageQuilting.jl:58-65
exports ``iqsim`` and ``voxelreuse``) plus one entry per hot-path function
(imfilter, fastdistance, band_pick, relaxation, taumodel, sample, graphcut).
All numerics run in libiqhip.so; there is no CPU fallback.
"""
from ._lib import IQError, LIB_PATH, EXPORTS  # noqa: F401
from .api import (  # noqa: F401
    Context,
    Plan,
    band_pick,
    default_context,
    device_count,
    fastdistance,
    finish,
    graphcut,
    imfilter,
    iqsim,
    iqsim_batch,
    make_plan,
    make_plan_multi,
    relaxation,
    sample,
    taumodel,
    voxelreuse,
)
from . import prep  # noqa: F401
