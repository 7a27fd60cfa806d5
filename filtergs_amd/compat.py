"""Import-name shims, so callers written against the reference run unmodified.

``install()`` registers this package's modules under the names the callers
import (LoG/render/renderer.py:1, FilterGS/filtergs.py:3-5):

* ``diff_gaussian_rasterization_wodilate`` (and its ``._C``) ->
  ``filtergs_amd.rasterizer`` / ``filtergs_amd.ext``;
* ``diff_gaussian_rasterization`` (and its ``._C``) -> ``filtergs_amd.dilate`` /
  ``filtergs_amd.dilate_ext`` (the 19-argument, 8-result graphdeco signatures);
* ``FilterGS.cuda.compute_radius``, ``FilterGS.cuda.visible_flag``,
  ``FilterGS.cuda.gather_attributes`` -> ``filtergs_amd.lod``;
* ``FilterGS`` / ``FilterGS.filtergs.FilterGS`` -> ``filtergs_amd.filtergs``
  when no FilterGS package is importable.

A ``FilterGS`` package that is importable keeps its own top-level module; only
its ``cuda`` subpackage is taken over, which also stops the reference's
just-in-time CUDA build at import (FilterGS/cuda/compute_radius.py:7-20).
Call ``install()`` before the first import of those names.
"""
from __future__ import annotations

import importlib.util
import sys
import types

from . import dilate, dilate_ext, ext, filtergs, lod, rasterizer

_FILTERGS_CUDA = ("compute_radius", "visible_flag", "gather_attributes")


def _alias(name, module):
    sys.modules[name] = module
    return module


def install(replace: bool = True) -> None:
    """Register the aliases (existing entries are replaced unless replace=False)."""
    def put(name, module):
        if replace or name not in sys.modules:
            _alias(name, module)

    put("diff_gaussian_rasterization_wodilate", rasterizer)
    put("diff_gaussian_rasterization_wodilate._C", ext)
    put("diff_gaussian_rasterization", dilate)
    put("diff_gaussian_rasterization._C", dilate_ext)

    if "FilterGS" not in sys.modules and importlib.util.find_spec("FilterGS") is None:
        # no FilterGS package here: provide this package's FilterGS class
        pkg = types.ModuleType("FilterGS")
        pkg.__path__ = []
        pkg.FilterGS = filtergs.FilterGS
        put("FilterGS", pkg)
        put("FilterGS.filtergs", filtergs)
    cuda = types.ModuleType("FilterGS.cuda")
    cuda.__path__ = []
    for sub in _FILTERGS_CUDA:
        setattr(cuda, sub, lod)
    put("FilterGS.cuda", cuda)
    for sub in _FILTERGS_CUDA:
        put(f"FilterGS.cuda.{sub}", lod)


def uninstall() -> None:
    """Drop the aliases that point at this package's modules."""
    ours = {rasterizer, dilate, dilate_ext, ext, lod, filtergs}
    for name in list(sys.modules):
        mod = sys.modules[name]
        if mod in ours and not name.startswith("filtergs_amd"):
            del sys.modules[name]
    for name in ("FilterGS.cuda", "FilterGS"):
        mod = sys.modules.get(name)
        if isinstance(mod, types.ModuleType) and getattr(mod, "__path__", None) == [] and \
                getattr(mod, "__file__", None) is None:
            del sys.modules[name]
