"""The FilterGS caller API (FilterGS/filtergs.py:9-359) on the gfx950 kernels.

``FilterGS`` keeps the reference's buffers (xyz, log-scale ``scaling``, raw
``rotation``, logit ``opacity``, SH DC ``colors``, SH rest ``shs``) and its three
stages:

* ``prepare`` -- the LoD cut of the tree (``compute_visibility_radius``), or
  without a tree the NDC frustum test plus the 3-sigma radius of the raw
  parameters (the reference computes that radius from the log-scales; see
  ``radius_from_log_scales``);
* ``get_all`` -- gather of the visible rows and their activation (exp, sigmoid,
  normalise, SH2RGB(dc) + SH without DC toward the camera) in one fused kernel,
  ``fgs_activate_gather``, instead of a gather plus five element-wise passes;
* ``render`` -- the finite / positive filter and the wodilate rasterizer with
  precomputed colours.  With this package's rasterizer, ``get_all`` and the
  filter are folded into the rasterizer's preprocess (``fgs_filtergs_geometry``
  / ``fgs_filtergs_render``): it reads row ``index[i]`` of the raw buffers and
  activates it while loading, so no gathered copy is made; other rasterizers
  get the gathered tensors as in the reference.
"""
from __future__ import annotations

import time

import torch
import torch.nn as nn

from . import _lib, ext, lod
from . import rasterizer as _rasterizer
from ._lib import call, ptr

SH_C0 = 0.28209479177387814


class FilterGS(nn.Module):
    def __init__(self, sh_degree=1, xyz_scale=1.):
        super().__init__()
        self.xyz_scale = xyz_scale
        self.max_sh_degree = sh_degree
        self.active_sh_degree = 0
        self.keys = ['xyz', 'features_dc', 'features_rest', 'scaling', 'rotation', 'opacity']
        self.visibility_flag = None
        self.sh_c0 = SH_C0
        # FilterGS/filtergs.py:261-278 feeds the raw log-scales to the no-tree
        # radius; keep that unless set to False
        self.radius_from_log_scales = True

    # -- construction (FilterGS/filtergs.py:304-359) -----------------------
    def register_by_pointcloud(self, xyz, colors, scales, init_opacity, **init_ply):
        """Initial buffers from a point cloud (the reference's
        register_by_pointcloud).  Each point gets an isotropic scale, clamped
        to [mean / 4, 4 mean] and stored as its log; its colour as SH DC,
        (c - 0.5) / C0; zero SH rest coefficients; the logit of init_opacity;
        the identity rotation.  With init_ply['height'] a ground grid is
        appended (create_from_ground); as in the reference only its scales go
        through the inverse activation (log), its colours (0.5) and opacities
        (ground_opacity) are stored as they are."""
        m = scales.mean()
        clamped = scales.clamp(min=m / 4, max=m * 4)
        print(f'[FilterGS] point-cloud scales {self.log_radius(scales)}, clamped to {self.log_radius(clamped)}')
        n = xyz.shape[0]
        o = torch.full((n, 1), init_opacity, dtype=xyz.dtype, device=xyz.device)
        cols = {'xyz': [xyz], 'colors': [(colors - 0.5) / SH_C0],
                'scaling': [clamped.log().unsqueeze(1).expand(n, 3).clone()],
                'opacity': [torch.log(o / (1 - o))], 'rotation': [self.init_rotation(n, xyz.device)]}
        rest = (self.max_sh_degree + 1) ** 2 - 1 if self.max_sh_degree > 0 else 0
        if rest:
            cols['shs'] = [torch.zeros((n, rest, 3), dtype=torch.float32)]
        if 'height' in init_ply:
            g_xyz, g_col, g_scale, g_op = self.create_from_ground(
                xyz.min(dim=0), xyz.max(dim=0), init_ply['init_step'], init_ply['height'],
                init_ply['ground_opacity'])
            k = g_xyz.shape[0]
            print(f'[FilterGS] {k} ground points appended')
            for name, t in (('xyz', g_xyz), ('colors', g_col), ('scaling', g_scale.log()), ('opacity', g_op),
                            ('rotation', self.init_rotation(k, xyz.device))):
                cols[name].append(t)
            if rest:
                cols['shs'].append(torch.zeros((k, rest, 3), dtype=torch.float32))
        for name in ('scaling', 'colors', 'xyz', 'opacity', 'rotation', 'shs'):
            if name in cols:
                self.register_buffer(name, torch.cat(cols[name], dim=0))
                self.keys.append(name)

    @staticmethod
    def create_from_ground(local_min, local_max, init_step, height, init_opacity=0.9, padding=0.05):
        """A flat ground grid at z = height: spacing init_step over the xy box
        of local_min / local_max (the (values, indices) results of
        xyz.min(dim=0) / xyz.max(dim=0)) widened by padding.  Returns the
        points, colours 0.5, scales (step, step, step / 10) and opacities
        init_opacity, none of them activated."""
        lo, hi = local_min[0], local_max[0]
        gx, gy = torch.meshgrid(torch.arange(lo[0] - padding, hi[0] + padding, init_step),
                                torch.arange(lo[1] - padding, hi[1] + padding, init_step), indexing="ij")
        n = gx.numel()
        xyz = torch.stack([gx.reshape(-1), gy.reshape(-1), torch.full((n,), float(height))], dim=1)
        scaling = torch.tensor([init_step, init_step, init_step * 0.1], dtype=torch.float32).repeat(n, 1)
        return xyz, torch.full((n, 3), 0.5), scaling, torch.full((n, 1), float(init_opacity))

    def init_rotation(self, num_points, device):
        """num_points identity quaternions (w, x, y, z) = (1, 0, 0, 0)."""
        return torch.tensor([1.0, 0.0, 0.0, 0.0], dtype=torch.float32, device=device).repeat(num_points, 1)

    def log_radius(self, scales):
        """min / mean / max of a scale tensor as text."""
        lo, mid, hi = (float(v) for v in (scales.min(), scales.mean(), scales.max()))
        return f'{lo:.4f} / {mid:.4f} / {hi:.4f}'

    # -- stage 1 -------------------------------------------------------------
    def prepare(self, rasterizer, camera, tree=None, tree_node_index_gpu=None, ancestor_path_gpu=None,
                radius_max=2.0, padding=0.5, radius_filter_threshold=700.0):
        """Visibility of the candidates: the LoD cut when a tree is given."""
        device = self.xyz.device
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record()
        if tree_node_index_gpu is not None and ancestor_path_gpu is not None:
            node_index_gpu = tree_node_index_gpu
        elif tree is not None and self.xyz.is_cuda:
            node_index_gpu = tree.node_index.to(device=device, dtype=torch.int32)
            ancestor_path_gpu = tree.ancestor_path.to(device=device, dtype=torch.int32)
        else:
            node_index_gpu, ancestor_path_gpu = None, None
        ev[1].record()
        all_mask, all_radius2d, all_depth = self._compute_visibility_radius_cuda(
            rasterizer, camera, node_index_gpu, ancestor_path_gpu, radius_max, padding,
            radius_filter_threshold)
        ev[2].record()
        index_all = torch.nonzero(all_mask, as_tuple=True)[0]
        radius2d = torch.masked_select(all_radius2d, all_mask)
        depth_visible = torch.masked_select(all_depth, all_mask) if all_depth.numel() > 0 else all_depth
        self.visibility_flag = {
            'index': index_all, 'flag': all_mask, 'radius2d': radius2d, 'depth': depth_visible,
            'raw_inputs': {
                'visible_index': index_all, 'node_index_gpu': node_index_gpu,
                'ancestor_path_gpu': ancestor_path_gpu,
                'base_attrs': {'xyz': self.xyz, 'scaling': self.scaling, 'rotation': self.rotation,
                               'opacity': self.opacity, 'colors': self.colors,
                               'shs': getattr(self, 'shs', None)}}}
        ev[3].record()
        ev[3].synchronize()
        self.last_prepare_timing = {
            "T_tree_ms": float(ev[0].elapsed_time(ev[1])),
            "T_vis_ms": float(ev[1].elapsed_time(ev[2])),
            "T_pack_ms": float(ev[2].elapsed_time(ev[3])),
            "T_prepare_total_ms": float(ev[0].elapsed_time(ev[3])),
            "vis_fallback": False,
            "num_visible": int(index_all.numel()),
        }
        return len(index_all)

    # -- stage 2 -------------------------------------------------------------
    def get_all(self, camera):
        """Gather + activation of the visible rows (one fused kernel)."""
        if self.visibility_flag is None:
            raise RuntimeError("Must call prepare() before get_all()")
        idx = self.visibility_flag['index']
        device = self.xyz.device
        _lib.require_gpu(self.xyz)
        idx = idx.to(device=device, dtype=torch.int64).contiguous()
        N = int(idx.numel())
        n_src = int(self.xyz.shape[0])
        f32 = dict(dtype=torch.float32, device=device)
        out = dict(xyz=torch.empty((N, 3), **f32), scaling=torch.empty((N, 3), **f32),
                   rotation=torch.empty((N, 4), **f32), opacity=torch.empty((N, 1), **f32),
                   colors=torch.empty((N, 3), **f32))
        deg = int(self.active_sh_degree)
        shs = getattr(self, 'shs', None)
        rest = shs.to(**f32).contiguous() if (deg > 0 and shs is not None) else None
        rest_coeffs = int(rest.shape[1]) if rest is not None else 0
        campos = None
        if deg > 0 and camera is not None:
            campos = torch.as_tensor(camera['camera_center']).to(**f32).reshape(3).contiguous()
        else:
            deg = 0
        bad = torch.zeros((1,), dtype=torch.int32, device=device)
        src = [t.to(**f32).contiguous() for t in (self.xyz, self.scaling, self.rotation,
                                                  self.opacity, self.colors)]
        if N > 0:
            call("fgs_activate_gather", N, idx.data_ptr(), n_src, *[ptr(t) for t in src], ptr(rest),
                 rest_coeffs, deg, ptr(campos), out['xyz'].data_ptr(), out['scaling'].data_ptr(),
                 out['rotation'].data_ptr(), out['opacity'].data_ptr(), out['colors'].data_ptr(),
                 bad.data_ptr(), _lib.stream_handle(device))
            if int(bad.item()):
                raise RuntimeError("gather_attributes: index out of range")
        return out

    # -- stage 3 -------------------------------------------------------------
    def render(self, camera, rasterizer, background=None, enable_stats=True):
        if self.visibility_flag is None:
            raise RuntimeError("Must call prepare() before render()")
        if isinstance(rasterizer, _rasterizer.GaussianRasterizer) and self.visibility_flag['index'].numel() > 0:
            return self._render_rows(camera, rasterizer, enable_stats)
        activated = self.get_all(camera)
        xyz, opacity, colors = activated['xyz'], activated['opacity'], activated['colors']
        scales, rotations = activated['scaling'], activated['rotation']
        if opacity.dim() == 1:
            opacity = opacity.unsqueeze(-1)
        finite = (torch.isfinite(xyz).all(dim=-1) & torch.isfinite(scales).all(dim=-1)
                  & torch.isfinite(rotations).all(dim=-1) & torch.isfinite(opacity).all(dim=-1)
                  & torch.isfinite(colors).all(dim=-1))
        finite = finite & (scales.min(dim=-1).values > 0.0) & (opacity[:, 0] > 0.0)
        if not torch.all(finite):
            xyz, opacity, colors = xyz[finite], opacity[finite], colors[finite]
            scales, rotations = scales[finite], rotations[finite]
        if xyz.shape[0] == 0:
            rs = rasterizer.raster_settings
            h, w = int(rs.image_height), int(rs.image_width)
            if background is None:
                bg = torch.zeros(3, dtype=xyz.dtype, device=xyz.device)
            else:
                bg = background.to(device=xyz.device, dtype=xyz.dtype).reshape(-1)
                if bg.numel() >= 3:
                    bg = bg[:3]
                elif bg.numel() == 1:
                    bg = bg.repeat(3)
                else:
                    bg = torch.zeros(3, dtype=xyz.dtype, device=xyz.device)
            return {"render": bg[:, None, None].expand(3, h, w).contiguous(),
                    "timing": {"T_raster_ms": 0.0}}
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        out = rasterizer(means3D=xyz, means2D=torch.zeros_like(xyz), shs=None, colors_precomp=colors,
                         opacities=opacity, scales=scales, rotations=rotations,
                         ogfs_enable_stats=bool(enable_stats))
        ev1.record()
        ev1.synchronize()
        return {"render": out[0][:3], "timing": {"T_raster_ms": float(ev0.elapsed_time(ev1))}}

    def _render_rows(self, camera, rasterizer, enable_stats):
        """get_all + render as one rasterizer pass over the visible rows: the
        preprocess reads and activates row index[i] itself, and rows the
        reference would drop as non-finite are culled there."""
        rs = rasterizer.raster_settings
        deg = int(self.active_sh_degree)
        campos = None
        if deg > 0 and camera is not None:
            campos = torch.as_tensor(camera['camera_center'])
        shs = getattr(self, 'shs', None)
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record()
        color, stats = ext.filtergs_render(
            rs, self.visibility_flag['index'], self.xyz, self.scaling, self.rotation, self.opacity,
            self.colors, shs if shs is not None else torch.Tensor([]), deg, campos,
            ogfs_epsilon_scale=rs.ogfs_epsilon_scale, ogfs_epsilon_max=rs.ogfs_epsilon_max,
            ogfs_energy_floor=rs.ogfs_energy_floor, ogfs_radius_ratio_min=rs.ogfs_radius_ratio_min,
            ogfs_enable_stats=bool(enable_stats))
        ev1.record()
        ev1.synchronize()
        _rasterizer._RasterizeGaussians._last_stats = stats
        rasterizer.latest_stats = _rasterizer.parse_stats(stats)
        return {"render": color[:3], "timing": {"T_raster_ms": float(ev0.elapsed_time(ev1))}}

    # -- visibility (FilterGS/filtergs.py:225-302) ---------------------------
    def _visible_flag_by_camera_cuda(self, xyz, camera, padding=0.05):
        return lod.visible_flag_cuda_direct(xyz, camera['full_proj_transform'], padding=padding)

    @staticmethod
    def _visible_flag_by_camera_torch(xyz, camera, padding=0.05):
        """The torch frustum test of the reference (NDC depth in (0, 1)),
        computed from the visible_flag kernel's projection."""
        _, _, p_proj = lod.visible_flag_cuda_direct(xyz, camera['full_proj_transform'], padding=padding)
        depth = p_proj[:, 2]
        valid = ((depth > 0.0) & (depth < 1.0) & (p_proj[:, 0] > -1.0 - padding)
                 & (p_proj[:, 0] < 1.0 + padding) & (p_proj[:, 1] > -1.0 - padding)
                 & (p_proj[:, 1] < 1.0 + padding))
        return valid, depth, p_proj

    def _compute_visibility_radius_cuda(self, rasterizer, camera, tree_node_index_gpu,
                                        ancestor_path_gpu, radius_max, padding,
                                        radius_filter_threshold=100.0):
        rs = rasterizer.raster_settings
        device = self.xyz.device
        tanfovx, tanfovy = float(rs.tanfovx), float(rs.tanfovy)
        focal_x = float(rs.image_width) / (2.0 * tanfovx)
        focal_y = float(rs.image_height) / (2.0 * tanfovy)
        proj = rs.projmatrix.to(device=device, dtype=torch.float32).contiguous()
        view = rs.viewmatrix.to(device=device, dtype=torch.float32).contiguous()
        scaling = self.scaling.detach().to(device=device).contiguous()
        rotation = self.rotation.detach().to(device=device).contiguous()
        if tree_node_index_gpu is None or ancestor_path_gpu is None:
            valid, depth, _ = self._visible_flag_by_camera_torch(self.xyz, camera, padding=padding)
            sc = scaling if self.radius_from_log_scales else scaling.exp()
            radius2d = lod.compute_radius_module.compute_radius(
                self.xyz.detach().contiguous(), sc, rotation, proj, view, focal_x, focal_y, tanfovx,
                tanfovy)
            mask = valid
            if radius_filter_threshold > 0:
                mask = mask & (radius2d <= radius_filter_threshold)
            return mask, radius2d, depth
        return lod.compute_visibility_radius(
            self.xyz.detach().contiguous(), scaling, rotation, proj, view, tree_node_index_gpu,
            ancestor_path_gpu, focal_x, focal_y, tanfovx, tanfovy, float(radius_max), padding,
            radius_filter_threshold)


__all__ = ["FilterGS"]
