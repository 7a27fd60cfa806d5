"""The wodilate plugin surface: GaussianRasterizationSettings,
GaussianRasterizer and the autograd function, with the argument handling,
defaults, outputs and errors of
antialias/diff_gaussian_rasterization_wodilate/__init__.py:12-345.

Added (keyword fields with defaults, so existing callers are unaffected):
``grad_mode`` -- "reference" reproduces the reference backward equations
(backward.cu, +0.3 on the covariance diagonal regardless of the forward
filter, alpha clamp passed through); "exact" differentiates the forward that
was actually computed.
"""
from __future__ import annotations

from typing import NamedTuple

import torch
import torch.nn as nn

from . import ext
from . import ext as _C  # noqa: F401 (the module's extension is reachable as _C)
from ._lib import guarded_call, host_copy as cpu_deep_copy_tuple  # noqa: F401 (module surface)

GRAD_MODES = {"reference": 0, "exact": 1}


def _grad_mode(settings) -> int:
    m = getattr(settings, "grad_mode", "reference")
    if isinstance(m, int):
        return m
    if m not in GRAD_MODES:
        raise ValueError(f"grad_mode must be one of {sorted(GRAD_MODES)}, got {m!r}")
    return GRAD_MODES[m]


def rasterize_gaussians(means3D, means2D, sh, colors_precomp, opacities, scales, rotations,
                        cov3Ds_precomp, raster_settings, use_filter=True, ogfs_epsilon_scale=0.2,
                        ogfs_epsilon_max=0.2, ogfs_energy_floor=0.05, ogfs_radius_ratio_min=0.7,
                        ogfs_enable_stats=True):
    return _RasterizeGaussians.apply(means3D, means2D, sh, colors_precomp, opacities, scales,
                                     rotations, cov3Ds_precomp, raster_settings, use_filter,
                                     float(ogfs_epsilon_scale), float(ogfs_epsilon_max),
                                     float(ogfs_energy_floor), float(ogfs_radius_ratio_min),
                                     bool(ogfs_enable_stats))


class _RasterizeGaussians(torch.autograd.Function):
    _last_stats = None
    variant = 0

    @staticmethod
    def forward(ctx, means3D, means2D, sh, colors_precomp, opacities, scales, rotations,
                cov3Ds_precomp, raster_settings, use_filter, ogfs_epsilon_scale, ogfs_epsilon_max,
                ogfs_energy_floor, ogfs_radius_ratio_min, ogfs_enable_stats):
        args = (raster_settings.bg, means3D, colors_precomp, opacities, scales, rotations,
                raster_settings.scale_modifier, cov3Ds_precomp, raster_settings.viewmatrix,
                raster_settings.projmatrix, raster_settings.tanfovx, raster_settings.tanfovy,
                raster_settings.image_height, raster_settings.image_width, sh,
                raster_settings.sh_degree, raster_settings.campos, raster_settings.prefiltered,
                use_filter, raster_settings.debug, ogfs_epsilon_scale, ogfs_epsilon_max,
                ogfs_energy_floor, ogfs_radius_ratio_min, ogfs_enable_stats)
        gm = _grad_mode(raster_settings)
        out = guarded_call(ext.rasterize_gaussians, args, dict(grad_mode=gm), raster_settings.debug,
                           "snapshot_fw.dump", "forward")
        (num_rendered, color, point_id, point_weight_pixel, point_weight, radii, geomBuffer,
         binningBuffer, imgBuffer, stats) = out
        _RasterizeGaussians._last_stats = stats
        ctx.raster_settings = raster_settings
        ctx.num_rendered = num_rendered
        ctx.use_filter = bool(use_filter)
        ctx.grad_mode = gm
        ctx.save_for_backward(colors_precomp, means3D, scales, rotations, cov3Ds_precomp, radii, sh,
                              geomBuffer, binningBuffer, imgBuffer, opacities)
        return color, radii, point_id, point_weight_pixel, point_weight

    @staticmethod
    def backward(ctx, grad_out_color, _radii, _point_id, _point_weight_pixel, _point_weight):
        num_rendered = ctx.num_rendered
        raster_settings = ctx.raster_settings
        (colors_precomp, means3D, scales, rotations, cov3Ds_precomp, radii, sh, geomBuffer,
         binningBuffer, imgBuffer, opacities) = ctx.saved_tensors
        args = (raster_settings.bg, means3D, radii, colors_precomp, scales, rotations,
                raster_settings.scale_modifier, cov3Ds_precomp, raster_settings.viewmatrix,
                raster_settings.projmatrix, raster_settings.tanfovx, raster_settings.tanfovy,
                grad_out_color, sh, raster_settings.sh_degree, raster_settings.campos, geomBuffer,
                num_rendered, binningBuffer, imgBuffer, raster_settings.debug)
        kw = dict(variant=_RasterizeGaussians.variant, use_filter=ctx.use_filter,
                  grad_mode=ctx.grad_mode, opacities=opacities)
        grads = guarded_call(ext.rasterize_gaussians_backward, args, kw, raster_settings.debug,
                             "snapshot_bw.dump", "backward")
        (grad_means2D, grad_colors_precomp, grad_opacities, grad_means3D, grad_cov3Ds_precomp,
         grad_sh, grad_scales, grad_rotations) = grads
        if grad_colors_precomp.shape[-1] > 3:
            grad_colors_precomp = grad_colors_precomp[:, :3]
        return (grad_means3D, grad_means2D, grad_sh, grad_colors_precomp,
                grad_opacities.view_as(opacities) if opacities.numel() == grad_opacities.numel() else grad_opacities,
                grad_scales, grad_rotations, grad_cov3Ds_precomp,
                None, None, None, None, None, None, None)


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
    ogfs_epsilon_scale: float = 0.2
    ogfs_epsilon_max: float = 0.2
    ogfs_energy_floor: float = 0.05
    ogfs_radius_ratio_min: float = 0.7
    ogfs_enable_stats: bool = True
    grad_mode: str = "reference"


def parse_stats(stats_tensor):
    """latest_stats dict from the packed stats tensor (wodilate/__init__.py:306-344)."""
    if stats_tensor is None or not torch.is_tensor(stats_tensor):
        return None
    stats_list = stats_tensor.detach().cpu().tolist()
    if len(stats_list) < 16:
        return None
    d = {
        "bar_g": float(stats_list[0]),
        "avg_kv_per_gaussian": float(stats_list[1]),
        "avg_kv_per_tile": float(stats_list[2]),
        "num_kv_pairs": float(stats_list[3]),
        "num_tiles": float(stats_list[4]),
        "num_gaussians": float(stats_list[5]),
    }
    num_kpc_bins = int(stats_list[8])
    d["used_kv_pairs"] = float(stats_list[6])
    d["unused_kv_pairs"] = float(stats_list[7])
    d["num_kpc_bins"] = num_kpc_bins
    remainder = stats_list[9:]
    num_tiles = int(d["num_tiles"])
    kpc_slots = 0
    if len(remainder) >= 2 * num_tiles:
        kpc_slots = (len(remainder) - 2 * num_tiles) // 2
    if kpc_slots > 0:
        kpc_part = remainder[:2 * kpc_slots]
        d["kpc_edges"] = kpc_part[:kpc_slots][:num_kpc_bins]
        d["kpc_hist"] = kpc_part[kpc_slots:][:num_kpc_bins]
        remainder = remainder[2 * kpc_slots:]
    if remainder and num_tiles > 0:
        if len(remainder) >= num_tiles:
            d["tile_ratio"] = remainder[:num_tiles]
            remainder = remainder[num_tiles:]
        if len(remainder) >= num_tiles:
            d["tile_total_kv"] = remainder[:num_tiles]
    return d


class GaussianRasterizer(nn.Module):
    def __init__(self, raster_settings):
        super().__init__()
        self.raster_settings = raster_settings
        self.latest_stats = None

    def compute_radius(self, positions, scales, rotations):
        with torch.no_grad():
            rs = self.raster_settings
            return ext.compute_radius(positions, scales, rotations, rs.viewmatrix, rs.projmatrix,
                                      rs.tanfovx, rs.tanfovy, rs.image_height, rs.image_width)

    def markVisible(self, positions):
        with torch.no_grad():
            rs = self.raster_settings
            return ext.mark_visible(positions, rs.viewmatrix, rs.projmatrix)

    def forward(self, means3D, means2D, opacities, shs=None, colors_precomp=None, scales=None,
                rotations=None, cov3D_precomp=None, use_filter=True, ogfs_epsilon_scale=None,
                ogfs_epsilon_max=None, ogfs_energy_floor=None, ogfs_radius_ratio_min=None,
                ogfs_enable_stats=None):
        rs = self.raster_settings
        eps_scale = rs.ogfs_epsilon_scale if ogfs_epsilon_scale is None else ogfs_epsilon_scale
        eps_max = rs.ogfs_epsilon_max if ogfs_epsilon_max is None else ogfs_epsilon_max
        energy_floor = rs.ogfs_energy_floor if ogfs_energy_floor is None else ogfs_energy_floor
        radius_min = rs.ogfs_radius_ratio_min if ogfs_radius_ratio_min is None else ogfs_radius_ratio_min
        enable_stats = rs.ogfs_enable_stats if ogfs_enable_stats is None else ogfs_enable_stats
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
        result = rasterize_gaussians(means3D, means2D, shs, colors_precomp, opacities, scales,
                                     rotations, cov3D_precomp, rs, use_filter, float(eps_scale),
                                     float(eps_max), float(energy_floor), float(radius_min),
                                     bool(enable_stats))
        self.latest_stats = parse_stats(getattr(_RasterizeGaussians, "_last_stats", None))
        return result
