"""The ``_C`` module of the dilate plugin surface.

``antialias/diff_gaussian_rasterization/__init__.py`` imports graphdeco's
un-vendored diff-gaussian-rasterization extension as ``_C`` and calls it with
the original, pre-OGFS signatures:

* ``rasterize_gaussians`` -- 19 positional arguments (bg, means3D, colors,
  opacities, scales, rotations, scale_modifier, cov3D_precomp, viewmatrix,
  projmatrix, tanfovx, tanfovy, image_height, image_width, sh, sh_degree,
  campos, prefiltered, debug), unpacked into 8 values (num_rendered, color,
  point_id, point_weight, radii, geomBuffer, binningBuffer, imgBuffer)
  (``__init__.py:60-92``);
* ``rasterize_gaussians_backward`` -- 21 positional arguments, 8 gradients
  (``__init__.py:103-143``);
* ``mark_visible`` (``__init__.py:171-181``).

Here they run variant 1 of this package's kernels: +0.3 on the diagonal of
the 2D covariance, 3-sigma tile radius, no OGFS, no statistics.
``point_weight`` is the per-Gaussian maximum of T*alpha, [P].  Keyword-only
extras: ``grad_mode`` (0 reference equations, 1 the exact derivative of the
forward) and ``opacities`` (needed by the backward in exact mode only).
"""
from __future__ import annotations

from . import ext
from ._lib import FGS_VARIANT_DILATE


def rasterize_gaussians(background, means3D, colors, opacity, scales, rotations, scale_modifier,
                        cov3D_precomp, viewmatrix, projmatrix, tan_fovx, tan_fovy, image_height,
                        image_width, sh, degree, campos, prefiltered, debug, *, grad_mode=0):
    out = ext.rasterize_gaussians(background, means3D, colors, opacity, scales, rotations,
                                  scale_modifier, cov3D_precomp, viewmatrix, projmatrix, tan_fovx,
                                  tan_fovy, image_height, image_width, sh, degree, campos,
                                  prefiltered, True, debug, ogfs_enable_stats=False,
                                  variant=FGS_VARIANT_DILATE, grad_mode=grad_mode)
    num_rendered, color, point_id, _weight_pixel, point_weight, radii, geom, binning, img, _ = out
    return num_rendered, color, point_id, point_weight, radii, geom, binning, img


def rasterize_gaussians_backward(background, means3D, radii, colors, scales, rotations,
                                 scale_modifier, cov3D_precomp, viewmatrix, projmatrix, tan_fovx,
                                 tan_fovy, dL_dout_color, sh, degree, campos, geomBuffer, R,
                                 binningBuffer, imageBuffer, debug, *, grad_mode=0, opacities=None):
    return ext.rasterize_gaussians_backward(
        background, means3D, radii, colors, scales, rotations, scale_modifier, cov3D_precomp,
        viewmatrix, projmatrix, tan_fovx, tan_fovy, dL_dout_color, sh, degree, campos, geomBuffer,
        R, binningBuffer, imageBuffer, debug, variant=FGS_VARIANT_DILATE, use_filter=True,
        grad_mode=grad_mode, opacities=opacities)


mark_visible = ext.mark_visible
