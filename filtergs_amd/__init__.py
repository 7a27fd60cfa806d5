"""MI355X-native (gfx950) FilterGS rasterizer and LoD-cut hot path.

Drop-in surfaces:
  filtergs_amd.rasterizer  -- diff_gaussian_rasterization_wodilate
  filtergs_amd.dilate      -- diff_gaussian_rasterization
  filtergs_amd.lod         -- FilterGS.cuda (LoD cut, radius, frustum, gather)
  filtergs_amd.ext         -- the _C extension functions
The kernels live in lib/libfiltergs_amd.so (C ABI: include/filtergs_amd.h).
"""
from .rasterizer import GaussianRasterizationSettings, GaussianRasterizer  # noqa: F401

__version__ = "0.1.0"
