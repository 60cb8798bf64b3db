"""MI355X-native 3-D Laplace FMM evaluation path (drop-in for minimal-ExaFMM 3d/).

The compute runs in ``libexafmm_mi355x.so`` (HIP, gfx950); :mod:`minimal_amd.fmm`
mirrors the reference's pass-level interface over its C ABI.
"""
from .inputs import BODY_DTYPE  # noqa: F401
