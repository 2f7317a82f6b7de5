"""MI355X DFTTest: the reference DFTTest() surface over a fused gfx950 HIP kernel."""

from .dfttest2 import Backend, DFTTest, DFTTest2, VideoClip, filter_clip, select_backend
from .params import FilterParams, dfttest2_params, dfttest_params, rdft

__all__ = [
    "Backend",
    "DFTTest",
    "DFTTest2",
    "FilterParams",
    "VideoClip",
    "dfttest2_params",
    "dfttest_params",
    "filter_clip",
    "rdft",
    "select_backend",
]
