"""Batched multi-view rasterization: V cameras over one set of Gaussians per
call, through the batched C ABI (fgs_batch_*).

The reference renders one camera per ``_C.rasterize_gaussians`` call
(antialias/rasterize_points.cu:36-260) and a training step loops over its
cameras.  ``BatchRasterizer`` renders all of them with one launch per stage:
the Gaussians are read once per step by the preprocess, one depth sort and one
tile sort cover every view, one blend launch renders every tile of every view,
and the backward folds the per-view gradients into one set of parameter
gradients.  The kv-pair buffer has a capacity, so a step needs no host read
and ``capture()`` can record it in a HIP graph; ``autotune()`` picks the blend
configuration (tile layouts, launch order or a persistent launch over a
cost-ordered (view, tile) queue, LDS batch depth) by timing the step.

Per view the forward outputs equal those of ``ext.rasterize_gaussians`` bit
for bit (images, radii, point ids, point weights); gradients equal the sum of
the per-view backward calls up to float summation order.

Per-view records and screen-space gradients live in the batch's element
layout: Gaussian i seen by view v is element v * P64 + 64 * (i // 64) + r,
with r the rank of i among the Gaussians of its 64-group that view v sees and
P64 = P rounded up to 64 (culled pairs have no element).  ``view_elements``,
``view_grads`` and ``view_records`` map them back to [P, ...] per view.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _lib
from ._lib import BatchSizes, BatchView, Gaussians, RasterSettings, call, ptr


def _mat16(m) -> list:
    a = m.detach().cpu().numpy() if torch.is_tensor(m) else np.asarray(m)
    return [float(x) for x in np.asarray(a, dtype=np.float32).reshape(-1)]


def _vec3(v) -> list:
    a = v.detach().cpu().numpy() if torch.is_tensor(v) else np.asarray(v)
    return [float(x) for x in np.asarray(a, dtype=np.float32).reshape(-1)[:3]]


class BatchRasterizer:
    """Forward and backward of V views of the same P Gaussians per call.

    ``cameras``: objects with ``width``, ``height``, ``tanfovx``, ``tanfovy``,
    ``world_view_transform`` and ``full_proj_transform`` (4x4, the transposed
    row-vector matrices of the plugin settings) and ``center`` (camera
    position).  ``bg``: one colour for every view or one per view.
    """

    def __init__(self, cameras: Sequence, P: int, *, bg=(0.0, 0.0, 0.0), sh_degree: int = 3,
                 variant: int = 0, use_filter: bool = True, scale_modifier: float = 1.0,
                 ogfs_epsilon_scale: float = 0.2, ogfs_epsilon_max: float = 0.2,
                 ogfs_energy_floor: float = 0.05, ogfs_radius_ratio_min: float = 0.7,
                 grad_mode: int = 0, debug: bool = False, device="cuda", k_headroom: float = 1.1):
        if not torch.cuda.is_available():
            raise _lib.FGSError("filtergs_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device(device)
        self.V = len(cameras)
        self.P = int(P)
        self.k_headroom = float(k_headroom)
        bgs = np.asarray(bg, dtype=np.float32).reshape(-1, 3)
        if bgs.shape[0] == 1:
            bgs = np.repeat(bgs, self.V, axis=0)
        views = (BatchView * self.V)()
        self.shapes = []
        for v, cam in enumerate(cameras):
            w = views[v]
            w.image_height = int(cam.height)
            w.image_width = int(cam.width)
            w.tanfovx = float(cam.tanfovx)
            w.tanfovy = float(cam.tanfovy)
            w.viewmatrix[:] = _mat16(cam.world_view_transform)
            w.projmatrix[:] = _mat16(cam.full_proj_transform)
            w.campos[:] = _vec3(cam.center)
            w.bg[:] = [float(x) for x in bgs[v]]
            self.shapes.append((int(cam.height), int(cam.width)))
        self._views = views
        lib = _lib.load()
        self.sizes = BatchSizes()
        call("fgs_batch_sizes_query", views, self.V, self.P, ctypes.byref(self.sizes))
        u8 = dict(dtype=torch.uint8, device=self.device)
        f32 = dict(dtype=torch.float32, device=self.device)
        self.plan = torch.empty((self.sizes.plan_bytes,), **u8)
        call("fgs_batch_plan", views, self.V, self.plan.data_ptr(), _lib.stream_handle(self.device))
        self.geom = torch.empty((self.sizes.geom_bytes,), **u8)
        self.image = torch.empty((self.sizes.image_bytes,), **u8)
        self.ws = torch.empty((self.sizes.backward_ws_bytes,), **u8)
        npix = int(self.sizes.pixels)
        self.radii = torch.zeros((self.V, self.P), dtype=torch.int32, device=self.device)
        self.color = torch.zeros((3 * npix,), **f32)
        self.point_id = torch.zeros((npix,), dtype=torch.int32, device=self.device)
        self.weight_pixel = torch.zeros((npix,), **f32)
        self.point_weight = torch.zeros((self.V, self.P), **f32)
        self.counts = torch.zeros((4,), dtype=torch.int64, device=self.device)
        self.P64 = (self.P + 63) // 64 * 64
        # per-view screen-space gradients in the element layout
        self.dL_dmeans2D = torch.zeros((self.V, self.P64, 3), **f32)
        self.dL_dcolors = torch.zeros((self.V, self.P64, 3), **f32)
        self.K_cap = 0
        self.binning = torch.empty((0,), **u8)
        self.pix_base = np.concatenate([[0], np.cumsum([h * w for h, w in self.shapes])]).astype(np.int64)
        s = RasterSettings()
        s.scale_modifier = float(scale_modifier)
        s.sh_degree = int(sh_degree)
        s.debug = int(bool(debug))
        s.variant = int(variant)
        s.use_filter = int(bool(use_filter))
        s.ogfs_epsilon_scale = float(ogfs_epsilon_scale)
        s.ogfs_epsilon_max = float(ogfs_epsilon_max)
        s.ogfs_energy_floor = float(ogfs_energy_floor)
        s.ogfs_radius_ratio_min = float(ogfs_radius_ratio_min)
        s.ogfs_enable_stats = 0
        s.grad_mode = int(grad_mode)
        self.settings = s
        self.sh_degree = int(sh_degree)
        self._g = None
        self._keep = None
        self._lib = lib

    # -- inputs ------------------------------------------------------------
    def bind(self, means3D, opacities, scales=None, rotations=None, shs=None, colors_precomp=None,
             cov3D_precomp=None):
        """Set the Gaussian tensors the next calls read (kept by reference:
        a captured graph reads whatever these tensors hold at replay)."""
        def f(t):
            if t is None or t.numel() == 0:
                return None
            if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
                raise ValueError("Gaussian tensors must be contiguous float32 on the GPU")
            return t
        keep = dict(means3D=f(means3D), opacities=f(opacities.reshape(-1)), scales=f(scales),
                    rotations=f(rotations), shs=f(shs), colors=f(colors_precomp),
                    cov3D=f(cov3D_precomp))
        if int(means3D.shape[0]) != self.P:
            raise ValueError("means3D has a different number of Gaussians than the plan")
        if keep["colors"] is not None and keep["colors"].numel() != self.V * self.P * 3:
            raise ValueError("colors_precomp of a batch is per view: [V, P, 3]")
        g = Gaussians()
        g.P = self.P
        g.M = int(shs.shape[1]) if keep["shs"] is not None else 0
        g.means3D = ptr(keep["means3D"])
        g.shs = ptr(keep["shs"])
        g.colors_precomp = ptr(keep["colors"])
        g.opacities = ptr(keep["opacities"])
        g.scales = ptr(keep["scales"])
        g.rotations = ptr(keep["rotations"])
        g.cov3D_precomp = ptr(keep["cov3D"])
        self._g, self._keep = g, keep

    def reserve(self, K_cap: int):
        """Size the kv-pair buffer for K_cap pairs."""
        K_cap = int(K_cap)
        if K_cap > self.K_cap:
            self.binning = torch.empty((self._lib.fgs_batch_binning_bytes(K_cap),), dtype=torch.uint8,
                                       device=self.device)
            self.K_cap = K_cap

    # -- passes ------------------------------------------------------------
    def _forward_launch(self):
        self.point_weight.zero_()
        call("fgs_batch_forward", ctypes.byref(self.settings), ctypes.byref(self._g),
             ctypes.byref(self.sizes), self.plan.data_ptr(), self.geom.data_ptr(),
             self.image.data_ptr(), ptr(self.binning), self.K_cap, self.radii.data_ptr(),
             self.color.data_ptr(), self.point_id.data_ptr(), self.weight_pixel.data_ptr(),
             self.point_weight.data_ptr(), self.counts.data_ptr(),
             _lib.stream_handle(self.device))

    def _backward_launch(self, dL_dpix, out):
        # dL_dmeans2D / dL_dcolors need no clearing: the backward clears the
        # visible elements and no other element is read
        call("fgs_batch_backward", ctypes.byref(self.settings), ctypes.byref(self._g),
             ctypes.byref(self.sizes), self.plan.data_ptr(), self.radii.data_ptr(),
             self.geom.data_ptr(), self.image.data_ptr(), ptr(self.binning), self.K_cap,
             dL_dpix.data_ptr(), self.dL_dmeans2D.data_ptr(), self.dL_dcolors.data_ptr(),
             out["opacities"].data_ptr(), out["means3D"].data_ptr(), out["cov3D"].data_ptr(),
             ptr(out["shs"]), ptr(out["scales"]), ptr(out["rotations"]), self.ws.data_ptr(),
             _lib.stream_handle(self.device))

    def forward(self):
        """Render every view.  Grows the kv-pair buffer when the views need
        more pairs than it holds (one host read of the pair count)."""
        if self._g is None:
            raise RuntimeError("bind() the Gaussians first")
        self._forward_launch()
        ch = self.counts.cpu()
        if int(ch[2]) & 1:
            raise RuntimeError("Point is filtered although prefiltered is set")
        K = int(ch[1])
        if K > self.K_cap:
            self.reserve(int(K * self.k_headroom) + 1024)
            self._forward_launch()
        return K

    def grad_buffers(self):
        """Output tensors of the parameter gradients (summed over views)."""
        f32 = dict(dtype=torch.float32, device=self.device)
        k = self._keep
        M = int(k["shs"].shape[1]) if k["shs"] is not None else 0
        return dict(means3D=torch.empty((self.P, 3), **f32), cov3D=torch.empty((self.P, 6), **f32),
                    opacities=torch.empty((self.P,), **f32),
                    shs=torch.empty((self.P, M, 3), **f32) if M else None,
                    scales=torch.empty((self.P, 3), **f32) if k["cov3D"] is None else None,
                    rotations=torch.empty((self.P, 4), **f32) if k["cov3D"] is None else None)

    def backward(self, dL_dpix, out=None):
        """dL_dpix: flat [3 * pixels] (the views' [3,H,W] blocks
        concatenated) or a list of [3,H,W] tensors.  Returns the parameter
        gradients summed over views; per-view dL/dmeans2D and dL/dcolors stay
        in ``dL_dmeans2D`` / ``dL_dcolors`` ([V, P64, 3], element layout; see
        ``view_grads``)."""
        if isinstance(dL_dpix, (list, tuple)):
            dL_dpix = torch.cat([t.reshape(-1) for t in dL_dpix])
        dL_dpix = dL_dpix.to(dtype=torch.float32).contiguous()
        if dL_dpix.numel() != 3 * int(self.sizes.pixels):
            raise ValueError("dL_dpix must hold 3 * pixels values")
        out = out if out is not None else self.grad_buffers()
        self._backward_launch(dL_dpix, out)
        return out

    # -- views of the outputs ---------------------------------------------
    def view_color(self, v):
        h, w = self.shapes[v]
        b = int(self.pix_base[v])
        return self.color[3 * b:3 * (b + h * w)].view(3, h, w)

    def view_pixels(self, t, v):
        h, w = self.shapes[v]
        b = int(self.pix_base[v])
        return t[b:b + h * w].view(h, w)

    def view_sorted(self, v, K):
        """View v's sorted kv list of the last forward, in the plain path's
        form: (tile [n] int64, gaussian [n] int64, ranges [T_v, 2] int64 into
        those arrays, (0, 0) for empty tiles).  The batch stores the pairs of
        all views sorted by (view-local tile, view, depth, index) with
        elements for values; this gathers view v's runs in tile order and
        maps the elements back to Gaussian indices.  K: the pair count the forward
        returned (at most the capacity)."""
        offs = (ctypes.c_size_t * 4)()
        call("fgs_batch_offsets", ctypes.byref(self.sizes), self.K_cap, offs)
        K = min(int(K), self.K_cap)
        T = int(self.sizes.tiles)
        rg = self.image[int(offs[1]):int(offs[1]) + 8 * T].view(torch.int32).view(T, 2)
        base = int(sum(((h + 15) // 16) * ((w + 15) // 16) for h, w in self.shapes[:v]))
        h, w = self.shapes[v]
        Tv = ((h + 15) // 16) * ((w + 15) // 16)
        r = rg[base:base + Tv].to(torch.int64)
        lens = (r[:, 1] - r[:, 0]).clamp_min(0)
        starts = torch.cumsum(lens, 0) - lens
        n = int(lens.sum())
        pos = torch.arange(n, device=r.device) - torch.repeat_interleave(starts, lens) + \
            torch.repeat_interleave(r[:, 0], lens)
        # tile keys (view-local tile << 6 | view) and elements, u32 arrays of
        # the capacity's length one after the other
        k0 = int(offs[0])
        keys = self.binning[k0:k0 + 4 * K].view(torch.int32).to(torch.int64) & 0xFFFFFFFF
        vals = self.binning[k0 + 4 * self.K_cap:k0 + 4 * self.K_cap + 4 * K].view(torch.int32)
        tile = keys[pos] >> 6
        elem = vals.to(torch.int64)[pos] & 0xFFFFFFFF
        ranges = torch.stack([starts, starts + lens], 1) * (lens > 0)[:, None]
        # elements back to Gaussian indices
        e = self.view_elements(v)
        inv = torch.zeros((self.P64,), dtype=torch.int64, device=self.device)
        m = e >= 0
        inv[e[m] - v * self.P64] = torch.arange(self.P, device=self.device, dtype=torch.int64)[m]
        return tile, inv[elem - v * self.P64], ranges

    def view_elements(self, v):
        """Element index (into [V * P64]) of each of view v's Gaussians after a
        forward, -1 where the view does not see it ([P] int64)."""
        vis = torch.zeros((self.P64,), dtype=torch.int64, device=self.device)
        vis[:self.P] = (self.radii[v] > 0).to(torch.int64)
        grp = vis.view(-1, 64)
        rank = torch.cumsum(grp, 1) - grp
        base = torch.arange(0, self.P64, 64, device=self.device, dtype=torch.int64)[:, None]
        e = (v * self.P64 + base + rank).view(-1)[:self.P]
        return torch.where(vis[:self.P] > 0, e, torch.full_like(e, -1))

    def _gather_rows(self, t, v, width):
        e = self.view_elements(v)
        flat = t.reshape(-1, width)
        out = torch.zeros((self.P, width), dtype=t.dtype, device=self.device)
        m = e >= 0
        out[m] = flat[e[m]]
        return out

    def view_grads(self, v):
        """dL/dmeans2D and dL/dcolors of view v after a backward, [P, 3]
        each (zero where the view does not see the Gaussian)."""
        return (self._gather_rows(self.dL_dmeans2D, v, 3), self._gather_rows(self.dL_dcolors, v, 3))

    def view_records(self, v):
        """The preprocess records of view v after a forward: means2D [P, 2],
        conic_opacity [P, 4], rgb [P, 3] (zero where the view does not see the
        Gaussian)."""
        from .ext import geom_views
        gv = geom_views(self.geom, self.V * self.P64)
        return {k: self._gather_rows(gv[k], v, w) for k, w in
                (("means2D", 2), ("conic_opacity", 4), ("rgb", 3))}

    def view_state(self, v):
        """final_T and n_contrib of view v ([H, W] each) after a forward."""
        offs = (ctypes.c_size_t * 4)()
        call("fgs_batch_offsets", ctypes.byref(self.sizes), self.K_cap, offs)
        npix = int(self.sizes.pixels)
        fT = self.image[int(offs[2]):int(offs[2]) + 4 * npix].view(torch.float32)
        nc = self.image[int(offs[3]):int(offs[3]) + 4 * npix].view(torch.int32)
        return self.view_pixels(fT, v), self.view_pixels(nc, v)

    # -- graph capture and autotuning ---------------------------------------
    def capture(self, dL_dpix, out=None, warmup: int = 2):
        """Record forward + backward in a HIP graph.  The capacity must
        already hold the views' pairs (run forward() first).  Returns
        (graph, grads); graph.replay() recomputes both from the bound tensors
        and dL_dpix.  After a replay, ``counts[1] <= K_cap`` must hold."""
        out = out if out is not None else self.grad_buffers()
        dL_dpix = dL_dpix.contiguous()
        side = torch.cuda.Stream(device=self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            for _ in range(warmup):
                self._forward_launch()
                self._backward_launch(dL_dpix, out)
        torch.cuda.current_stream(self.device).wait_stream(side)
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            self._forward_launch()
            self._backward_launch(dL_dpix, out)
        return graph, out

    # Blend configurations the autotuner tries: each entry overrides the
    # defaults in _TUNE_BASE.  "tile_queue" > 0 runs the one-wave blends as
    # persistent launches over the cost-ordered (view, tile) queue with that
    # many tiles per XCD run; "xcd_chunk" is the run length of the launch
    # order without the queue; the layouts pick one wave or four waves per
    # tile; "work_shape" is the four-wave layout's LDS batch depth.
    _TUNE_BASE = {"forward_layout": 2, "backward_layout": 1, "xcd_chunk": 32, "tile_queue": 0,
                  "by_cost": True, "work_shape": (256, 128)}
    TUNE_CONFIGS = {
        "launch_xcd32": {},
        "launch_xcd0": {"xcd_chunk": 0},
        "queue8": {"tile_queue": 8},
        "queue32": {"tile_queue": 32},
        "queue128": {"tile_queue": 128},
        "queue32_launch_order": {"tile_queue": 32, "by_cost": False},
        "queue1_launch_order": {"tile_queue": 1, "by_cost": False},
        "fwd4wave_256": {"forward_layout": 0},
        "bwd4wave_128": {"backward_layout": 0},
    }

    @staticmethod
    def apply_config(cfg):
        """Set the process-wide blend knobs of one autotuner configuration."""
        c = dict(BatchRasterizer._TUNE_BASE)
        c.update(cfg)
        _lib.set_forward_layout(c["forward_layout"])
        _lib.set_backward_layout(c["backward_layout"])
        _lib.set_xcd_chunk(c["xcd_chunk"])
        _lib.set_tile_queue(c["tile_queue"], c["by_cost"])
        _lib.set_work_shape(*c["work_shape"])

    def autotune(self, dL_dpix, configs=None, iters: int = 3):
        """Time forward + backward under each blend configuration (layouts,
        tile order: launch order or persistent queue and its run length, LDS
        batch depth) and keep the fastest, process-wide.  The outputs do not
        depend on the choice.  Returns {label: ms}."""
        configs = self.TUNE_CONFIGS if configs is None else configs
        out = self.grad_buffers()
        res = {}
        for label, cfg in configs.items():
            self.apply_config(cfg)
            self._forward_launch()
            self._backward_launch(dL_dpix, out)
            t0 = torch.cuda.Event(enable_timing=True)
            t1 = torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(iters):
                self._forward_launch()
                self._backward_launch(dL_dpix, out)
            t1.record()
            t1.synchronize()
            res[label] = t0.elapsed_time(t1) / iters
        best = min(res, key=res.get)
        self.apply_config(configs[best])
        self.tuned = best
        return res
