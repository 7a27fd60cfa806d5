"""The dilate plugin surface (antialias/diff_gaussian_rasterization/__init__.py:12-222):
+0.3 low-pass dilation of the 2D covariance, 3-sigma tile radius, raw opacity,
no OGFS.  Its ``_C`` is graphdeco's un-vendored diff-gaussian-rasterization;
the forward follows that published algorithm (the +0.3 is pinned by
antialias/cuda_rasterizer/backward.cu:197-199 and LoG/model/geometry.py:87-88)
and returns (color, radii, point_id, point_weight) with point_weight the
per-Gaussian maximum of T*alpha, [P].
"""
from __future__ import annotations

from typing import NamedTuple

import torch
import torch.nn as nn

from . import dilate_ext as _C
from ._lib import guarded_call, host_copy as cpu_deep_copy_tuple  # noqa: F401 (module surface)
from .rasterizer import GRAD_MODES


def rasterize_gaussians(means3D, means2D, sh, colors_precomp, opacities, scales, rotations,
                        cov3Ds_precomp, raster_settings):
    return _RasterizeGaussians.apply(means3D, means2D, sh, colors_precomp, opacities, scales,
                                     rotations, cov3Ds_precomp, raster_settings)


def _gm(rs):
    return GRAD_MODES[getattr(rs, "grad_mode", "reference")]


class _RasterizeGaussians(torch.autograd.Function):
    """Forward and backward through ``_C`` with the dilate module's own
    19- and 21-argument calls (``__init__.py:60-143``)."""

    @staticmethod
    def forward(ctx, means3D, means2D, sh, colors_precomp, opacities, scales, rotations,
                cov3Ds_precomp, raster_settings):
        rs = raster_settings
        args = (rs.bg, means3D, colors_precomp, opacities, scales, rotations, rs.scale_modifier,
                cov3Ds_precomp, rs.viewmatrix, rs.projmatrix, rs.tanfovx, rs.tanfovy,
                rs.image_height, rs.image_width, sh, rs.sh_degree, rs.campos, rs.prefiltered,
                rs.debug)
        (num_rendered, color, point_id, point_weight, radii, geomBuffer, binningBuffer,
         imgBuffer) = guarded_call(_C.rasterize_gaussians, args, dict(grad_mode=_gm(rs)), rs.debug,
                                   "snapshot_fw.dump", "forward")
        ctx.raster_settings = rs
        ctx.num_rendered = num_rendered
        ctx.save_for_backward(colors_precomp, means3D, scales, rotations, cov3Ds_precomp, radii, sh,
                              geomBuffer, binningBuffer, imgBuffer, opacities)
        return color, radii, point_id, point_weight

    @staticmethod
    def backward(ctx, grad_out_color, _radii, _point_id, _point_weight):
        rs = ctx.raster_settings
        (colors_precomp, means3D, scales, rotations, cov3Ds_precomp, radii, sh, geomBuffer,
         binningBuffer, imgBuffer, opacities) = ctx.saved_tensors
        args = (rs.bg, means3D, radii, colors_precomp, scales, rotations, rs.scale_modifier,
                cov3Ds_precomp, rs.viewmatrix, rs.projmatrix, rs.tanfovx, rs.tanfovy,
                grad_out_color, sh, rs.sh_degree, rs.campos, geomBuffer, ctx.num_rendered,
                binningBuffer, imgBuffer, rs.debug)
        grads = guarded_call(_C.rasterize_gaussians_backward, args,
                             dict(grad_mode=_gm(rs), opacities=opacities), rs.debug,
                             "snapshot_bw.dump", "backward")
        (grad_means2D, grad_colors_precomp, grad_opacities, grad_means3D, grad_cov3Ds_precomp,
         grad_sh, grad_scales, grad_rotations) = grads
        if grad_colors_precomp.shape[-1] > 3:
            grad_colors_precomp = grad_colors_precomp[:, :3]
        if opacities.numel() == grad_opacities.numel():
            grad_opacities = grad_opacities.view_as(opacities)
        return (grad_means3D, grad_means2D, grad_sh, grad_colors_precomp, grad_opacities,
                grad_scales, grad_rotations, grad_cov3Ds_precomp, None)


class GaussianRasterizationSettings(NamedTuple):
    image_height: int
    image_width: int
    tanfovx: float
    tanfovy: float
    bg: torch.Tensor
    scale_modifier: float
    viewmatrix: torch.Tensor
    projmatrix: torch.Tensor
    sh_degree: int
    campos: torch.Tensor
    prefiltered: bool
    debug: bool
    grad_mode: str = "reference"


class GaussianRasterizer(nn.Module):
    def __init__(self, raster_settings):
        super().__init__()
        self.raster_settings = raster_settings

    def markVisible(self, positions):
        with torch.no_grad():
            rs = self.raster_settings
            return _C.mark_visible(positions, rs.viewmatrix, rs.projmatrix)

    def forward(self, means3D, means2D, opacities, shs=None, colors_precomp=None, scales=None,
                rotations=None, cov3D_precomp=None):
        rs = self.raster_settings
        if (shs is None and colors_precomp is None) or (shs is not None and colors_precomp is not None):
            raise Exception('Please provide excatly one of either SHs or precomputed colors!')
        if ((scales is None or rotations is None) and cov3D_precomp is None) or \
                ((scales is not None or rotations is not None) and cov3D_precomp is not None):
            raise Exception('Please provide exactly one of either scale/rotation pair or precomputed 3D covariance!')
        if shs is None:
            shs = torch.Tensor([])
        if colors_precomp is None:
            colors_precomp = torch.Tensor([])
        if scales is None:
            scales = torch.Tensor([])
        if rotations is None:
            rotations = torch.Tensor([])
        if cov3D_precomp is None:
            cov3D_precomp = torch.Tensor([])
        return rasterize_gaussians(means3D, means2D, shs, colors_precomp, opacities, scales,
                                   rotations, cov3D_precomp, rs)
