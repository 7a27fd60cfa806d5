"""Functions with the call signatures of the reference's ``_C`` extension
(antialias/ext.cpp:17-50, antialias/rasterize_points.cu:36-318), running the
gfx950 kernels through the C ABI.

``rasterize_gaussians`` returns the same 10-tuple (num_rendered, color,
point_id, point_weight_pixel, point_weight, radii, geomBuffer, binningBuffer,
imgBuffer, stats) and ``rasterize_gaussians_backward`` the same 8-tuple.
Keyword-only extras select the dilate variant and the gradient mode.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import Gaussians, RasterSettings, call, ptr

MAX_KPC_BINS = 64


def _nonempty(t):
    return t is not None and torch.is_tensor(t) and t.numel() > 0


def _f32(t, device):
    if not _nonempty(t):
        return None
    if t.dtype != torch.float32:
        raise RuntimeError(f"expected a float32 tensor, got {t.dtype}")
    return t.to(device=device).contiguous()


def _settings(bg, viewmatrix, projmatrix, campos, image_height, image_width, tanfovx, tanfovy,
              scale_modifier, sh_degree, prefiltered, debug, variant, use_filter, eps_scale,
              eps_max, energy_floor, ratio_min, enable_stats, grad_mode, device):
    keep = [t.to(device=device, dtype=torch.float32).contiguous()
            for t in (bg, viewmatrix, projmatrix, campos)]
    s = RasterSettings()
    s.image_height = int(image_height)
    s.image_width = int(image_width)
    s.tanfovx = float(tanfovx)
    s.tanfovy = float(tanfovy)
    s.bg = keep[0].data_ptr()
    s.scale_modifier = float(scale_modifier)
    s.viewmatrix = keep[1].data_ptr()
    s.projmatrix = keep[2].data_ptr()
    s.sh_degree = int(sh_degree)
    s.campos = keep[3].data_ptr()
    s.prefiltered = int(bool(prefiltered))
    s.debug = int(bool(debug))
    s.variant = int(variant)
    s.use_filter = int(bool(use_filter))
    s.ogfs_epsilon_scale = float(eps_scale)
    s.ogfs_epsilon_max = float(eps_max)
    s.ogfs_energy_floor = float(energy_floor)
    s.ogfs_radius_ratio_min = float(ratio_min)
    s.ogfs_enable_stats = int(bool(enable_stats))
    s.grad_mode = int(grad_mode)
    return s, keep


def _gaussians(P, means3D, sh, colors, opacity, scales, rotations, cov3D_precomp, device):
    t = dict(means3D=_f32(means3D, device), shs=_f32(sh, device), colors=_f32(colors, device),
             opacity=_f32(opacity, device), scales=_f32(scales, device),
             rotations=_f32(rotations, device), cov3D=_f32(cov3D_precomp, device))
    g = Gaussians()
    g.P = int(P)
    g.M = int(sh.size(1)) if _nonempty(sh) else 0
    g.means3D = ptr(t["means3D"])
    g.shs = ptr(t["shs"])
    g.colors_precomp = ptr(t["colors"])
    g.opacities = ptr(t["opacity"])
    g.scales = ptr(t["scales"])
    g.rotations = ptr(t["rotations"])
    g.cov3D_precomp = ptr(t["cov3D"])
    return g, t


def _kpc_used_thresh():
    """KPC_USED_THRESH as rasterizer_impl.cu:44-55 reads it (default 1e-4)."""
    import os
    env = os.environ.get("KPC_USED_THRESH", "")
    if env:
        try:
            v = float(env)
            return v if v > 0 else 1e-4
        except ValueError:
            return 1e-4
    return 1e-4


def _stats_head(P, T, K):
    """RenderStats of a forward without kv pairs (rasterize_points.cu:138-177)."""
    head = torch.zeros(9 + 2 * MAX_KPC_BINS, dtype=torch.float64)
    head[4] = float(T)
    head[5] = float(P)
    head[3] = float(K)
    if P > 0:
        head[1] = float(K) / float(P)
    if T > 0:
        head[2] = float(K) / float(T)
    return head


def rasterize_gaussians(background, means3D, colors, opacity, scales, rotations, scale_modifier,
                        cov3D_precomp, viewmatrix, projmatrix, tan_fovx, tan_fovy, image_height,
                        image_width, sh, degree, campos, prefiltered, use_filter, debug=False,
                        ogfs_epsilon_scale=0.2, ogfs_epsilon_max=0.2, ogfs_energy_floor=0.05,
                        ogfs_radius_ratio_min=0.7, ogfs_enable_stats=True, *, variant=0,
                        grad_mode=0):
    """RasterizeGaussiansCUDA (antialias/rasterize_points.cu:36-179)."""
    if means3D.ndim != 2 or means3D.size(1) != 3:
        raise RuntimeError("means3D must have dimensions (num_points, 3)")
    _lib.require_gpu(means3D)
    dev = means3D.device
    P = int(means3D.size(0))
    H, W = int(image_height), int(image_width)
    gx, gy = (W + 15) // 16, (H + 15) // 16
    T = gx * gy
    fopt = dict(dtype=torch.float32, device=dev)
    out_color = torch.zeros((3, H, W), **fopt)
    out_point_id = torch.zeros((H, W), dtype=torch.int32, device=dev)
    out_wpix = torch.zeros((H, W), **fopt)
    radii = torch.zeros((P,), dtype=torch.int32, device=dev)
    out_pw = torch.zeros((P,), **fopt)
    u8 = dict(dtype=torch.uint8, device=dev)
    geom = torch.empty((0,), **u8)
    binning = torch.empty((0,), **u8)
    img = torch.empty((0,), **u8)
    K = 0
    if P != 0:
        s, keep_s = _settings(background, viewmatrix, projmatrix, campos, H, W, tan_fovx, tan_fovy,
                              scale_modifier, degree, prefiltered, debug, variant, use_filter,
                              ogfs_epsilon_scale, ogfs_epsilon_max, ogfs_energy_floor,
                              ogfs_radius_ratio_min, ogfs_enable_stats, grad_mode, dev)
        g, keep_g = _gaussians(P, means3D, sh, colors, opacity, scales, rotations, cov3D_precomp, dev)
        lib = _lib.load()
        geom = torch.empty((lib.fgs_geom_bytes(P),), **u8)
        img = torch.empty((lib.fgs_image_bytes(W, H),), **u8)
        counts = torch.empty((4,), dtype=torch.int64, device=dev)
        stream = _lib.stream_handle(dev)
        call("fgs_forward_geometry", ctypes.byref(s), ctypes.byref(g), geom.data_ptr(),
             radii.data_ptr(), counts.data_ptr(), stream)
        ch = counts.cpu()   # the reference's num_rendered readback (rasterizer_impl.cu:367)
        if int(ch[2]) & 1:
            raise RuntimeError("Point is filtered although prefiltered is set")
        K = int(ch[1])
        if K > 0x7FFFFFFF:
            raise RuntimeError(f"too many kv pairs ({K})")
        binning = torch.empty((lib.fgs_binning_bytes(K, W, H),), **u8)
        collect = bool(ogfs_enable_stats) and K > 0
        if collect:
            stats_dev = torch.empty((lib.fgs_stats_words(W, H),), dtype=torch.float64, device=dev)
        call("fgs_forward_render_stats", ctypes.byref(s), ctypes.byref(g), radii.data_ptr(),
             geom.data_ptr(), K, binning.data_ptr(), img.data_ptr(), out_color.data_ptr(),
             out_point_id.data_ptr(), out_wpix.data_ptr(), out_pw.data_ptr(), None, None,
             stats_dev.data_ptr() if collect else None, _kpc_used_thresh(), stream)
        del keep_s, keep_g
        if collect:
            return (K, out_color, out_point_id, out_wpix, out_pw, radii, geom, binning, img,
                    stats_dev[:9 + 2 * MAX_KPC_BINS + 2 * T].cpu())
    stats = _stats_head(P, T, K)
    return (K, out_color, out_point_id, out_wpix, out_pw, radii, geom, binning, img, stats)


def filtergs_render(raster_settings, index, xyz, scaling, rotation, opacity, colors, shs,
                    sh_degree, campos, use_filter=True, ogfs_epsilon_scale=0.2, ogfs_epsilon_max=0.2,
                    ogfs_energy_floor=0.05, ogfs_radius_ratio_min=0.7, ogfs_enable_stats=True):
    """FilterGS.get_all + FilterGS.render in one rasterizer pass
    (FilterGS/filtergs.py:98-223): the preprocess reads row index[i] of the
    raw parameters and activates it (no gathered copy, no finite filter pass).
    Returns (color [3,H,W], stats) like the rasterizer call of render()."""
    rs = raster_settings
    dev = xyz.device
    _lib.require_gpu(xyz)
    H, W = int(rs.image_height), int(rs.image_width)
    T = ((W + 15) // 16) * ((H + 15) // 16)
    idx = index.to(device=dev, dtype=torch.int64).contiguous()
    N = int(idx.numel())
    f32 = dict(dtype=torch.float32, device=dev)
    src = [t.to(**f32).contiguous() for t in (xyz, scaling, rotation, opacity, colors)]
    deg = int(sh_degree)
    rest = shs.to(**f32).contiguous() if (deg > 0 and _nonempty(shs)) else None
    cp = campos.to(**f32).reshape(3).contiguous() if (deg > 0 and campos is not None) else None
    if rest is None or cp is None:
        deg = 0
    out_color = torch.zeros((3, H, W), **f32)
    s, keep_s = _settings(rs.bg, rs.viewmatrix, rs.projmatrix, rs.campos, H, W, rs.tanfovx, rs.tanfovy,
                          rs.scale_modifier, 0, rs.prefiltered, rs.debug, 0, use_filter,
                          ogfs_epsilon_scale, ogfs_epsilon_max, ogfs_energy_floor,
                          ogfs_radius_ratio_min, ogfs_enable_stats, 0, dev)
    f = _lib.FilterGSRows()
    f.N = N
    f.index = idx.data_ptr()
    f.n_src = int(src[0].shape[0])
    f.xyz, f.scaling, f.rotation, f.opacity, f.colors = [t.data_ptr() for t in src]
    f.shs = ptr(rest)
    f.rest_coeffs = int(rest.shape[1]) if rest is not None else 0
    f.sh_degree = deg
    f.campos = ptr(cp)
    lib = _lib.load()
    stream = _lib.stream_handle(dev)
    u8 = dict(dtype=torch.uint8, device=dev)
    radii = torch.zeros((N,), dtype=torch.int32, device=dev)
    geom = torch.empty((lib.fgs_geom_bytes(N),), **u8)
    img = torch.empty((lib.fgs_image_bytes(W, H),), **u8)
    counts = torch.empty((4,), dtype=torch.int64, device=dev)
    call("fgs_filtergs_geometry", ctypes.byref(s), ctypes.byref(f), geom.data_ptr(), radii.data_ptr(),
         counts.data_ptr(), stream)
    ch = counts.cpu()
    if int(ch[2]) & 2:
        raise RuntimeError("gather_attributes: index out of range")
    K = int(ch[1])
    if K > 0x7FFFFFFF:
        raise RuntimeError(f"too many kv pairs ({K})")
    binning = torch.empty((lib.fgs_binning_bytes(K, W, H),), **u8)
    collect = bool(ogfs_enable_stats) and K > 0
    stats_dev = torch.empty((lib.fgs_stats_words(W, H),), dtype=torch.float64, device=dev) if collect else None
    point_id = torch.empty((H, W), dtype=torch.int32, device=dev)
    wpix = torch.empty((H, W), **f32)
    pw = torch.zeros((N,), **f32)
    call("fgs_filtergs_render", ctypes.byref(s), ctypes.byref(f), radii.data_ptr(), geom.data_ptr(), K,
         binning.data_ptr(), img.data_ptr(), out_color.data_ptr(), point_id.data_ptr(), wpix.data_ptr(),
         pw.data_ptr(), ptr(stats_dev), _kpc_used_thresh(), stream)
    del keep_s
    if collect:
        return out_color, stats_dev[:9 + 2 * MAX_KPC_BINS + 2 * T].cpu()
    return out_color, _stats_head(N, T, K)


def rasterize_gaussians_backward(background, means3D, radii, colors, scales, rotations,
                                 scale_modifier, cov3D_precomp, viewmatrix, projmatrix, tan_fovx,
                                 tan_fovy, dL_dout_color, sh, degree, campos, geomBuffer, R,
                                 binningBuffer, imageBuffer, debug=False, *, variant=0,
                                 use_filter=True, grad_mode=0, opacities=None):
    """RasterizeGaussiansBackwardCUDA (antialias/rasterize_points.cu:181-260)."""
    dev = means3D.device
    P = int(means3D.size(0))
    H, W = int(dL_dout_color.size(1)), int(dL_dout_color.size(2))
    M = int(sh.size(1)) if _nonempty(sh) else 0
    fopt = dict(dtype=torch.float32, device=dev)
    dL_dmeans2D = torch.zeros((P, 3), **fopt)
    dL_dcolors = torch.zeros((P, 3), **fopt)
    dL_dopacity = torch.zeros((P, 1), **fopt)
    has_sr = _nonempty(scales) and not _nonempty(cov3D_precomp)
    alloc = torch.empty if P > 0 else torch.zeros
    dL_dmeans3D = alloc((P, 3), **fopt)
    dL_dcov3D = alloc((P, 6), **fopt)
    dL_dsh = (alloc if (M > 0 and not _nonempty(colors)) else torch.zeros)((P, M, 3), **fopt)
    dL_dscales = (alloc if has_sr else torch.zeros)((P, 3), **fopt)
    dL_drot = (alloc if has_sr else torch.zeros)((P, 4), **fopt)
    if P != 0:
        _lib.require_gpu(means3D, dL_dout_color)
        s, keep_s = _settings(background, viewmatrix, projmatrix, campos, H, W, tan_fovx, tan_fovy,
                              scale_modifier, degree, False, debug, variant, use_filter, 0.2, 0.2,
                              0.05, 0.7, False, grad_mode, dev)
        if opacities is None:
            if int(grad_mode) == 1:
                raise RuntimeError("grad_mode 'exact' needs the opacities in the backward")
            opacities = torch.ones((P,), **fopt)   # never read by the reference equations
        g, keep_g = _gaussians(P, means3D, sh, colors, opacities.reshape(-1), scales, rotations,
                               cov3D_precomp, dev)
        lib = _lib.load()
        ws = torch.empty((lib.fgs_backward_ws_bytes(P),), dtype=torch.uint8, device=dev)
        dpix = dL_dout_color.to(dtype=torch.float32).contiguous()
        call("fgs_backward", ctypes.byref(s), ctypes.byref(g), radii.contiguous().data_ptr(),
             geomBuffer.data_ptr(), int(R), ptr(binningBuffer), imageBuffer.data_ptr(),
             dpix.data_ptr(), dL_dmeans2D.data_ptr(), dL_dcolors.data_ptr(), dL_dopacity.data_ptr(),
             dL_dmeans3D.data_ptr(), dL_dcov3D.data_ptr(), ptr(dL_dsh), ptr(dL_dscales),
             ptr(dL_drot), ws.data_ptr(), _lib.stream_handle(dev))
        del keep_s, keep_g
    return (dL_dmeans2D, dL_dcolors, dL_dopacity, dL_dmeans3D, dL_dcov3D, dL_dsh, dL_dscales,
            dL_drot)


def compute_radius(means3D, scales, rotations, viewmatrix, projmatrix, tan_fovx, tan_fovy,
                   image_height, image_width):
    """compute_radius (antialias/rasterize_points.cu:263-296): float radii [P]."""
    P = int(means3D.size(0))
    radii = torch.zeros((P,), dtype=means3D.dtype, device=means3D.device)
    if P != 0:
        _lib.require_gpu(means3D)
        dev = means3D.device
        m, sc, ro = (_f32(t, dev) for t in (means3D, scales, rotations))
        vm = viewmatrix.to(device=dev, dtype=torch.float32).contiguous()
        pm = projmatrix.to(device=dev, dtype=torch.float32).contiguous()
        call("fgs_compute_radius", P, ptr(m), ptr(sc), ptr(ro), ptr(vm), ptr(pm), float(tan_fovx),
             float(tan_fovy), int(image_height), int(image_width), radii.data_ptr(),
             _lib.stream_handle(dev))
    return radii


def mark_visible(means3D, viewmatrix, projmatrix):
    """markVisible (antialias/rasterize_points.cu:299-318): bool [P]."""
    P = int(means3D.size(0))
    present = torch.zeros((P,), dtype=torch.bool, device=means3D.device)
    if P != 0:
        _lib.require_gpu(means3D)
        dev = means3D.device
        m = _f32(means3D, dev)
        vm = viewmatrix.to(device=dev, dtype=torch.float32).contiguous()
        pm = projmatrix.to(device=dev, dtype=torch.float32).contiguous()
        call("fgs_mark_visible", P, ptr(m), ptr(vm), ptr(pm), present.data_ptr(),
             _lib.stream_handle(dev))
    return present


def read_sorted(geomBuffer, P, num_rendered, binningBuffer):
    """The sorted kv list of the last forward: (keys u64 as int64, values int32)."""
    K = int(num_rendered)
    dev = geomBuffer.device
    keys = torch.empty((K,), dtype=torch.int64, device=dev)
    vals = torch.empty((K,), dtype=torch.int32, device=dev)
    if K > 0:
        call("fgs_read_sorted", geomBuffer.data_ptr(), int(P), K, binningBuffer.data_ptr(),
             keys.data_ptr(), vals.data_ptr(), _lib.stream_handle(dev))
    return keys, vals


def image_ranges(imageBuffer, image_width, image_height):
    """Tile ranges [T,2] (int32 view into the image buffer)."""
    W, H = int(image_width), int(image_height)
    T = ((W + 15) // 16) * ((H + 15) // 16)
    rp = _lib.load().fgs_image_ranges(imageBuffer.data_ptr(), W, H)
    off = rp - imageBuffer.data_ptr()
    return imageBuffer[off:off + 8 * T].view(torch.int32).view(T, 2)


def image_views(imageBuffer, image_width, image_height):
    """Views of the per-pixel state inside an image buffer:
    dict(final_T [H,W] float32, n_contrib [H,W] int32, ranges [T,2] int32)."""
    W, H = int(image_width), int(image_height)
    T = ((W + 15) // 16) * ((H + 15) // 16)
    offs = (ctypes.c_size_t * 3)()
    _lib.load().fgs_image_offsets(W, H, offs)

    def view(i, n, dtype, shape):
        o = int(offs[i])
        return imageBuffer[o:o + 4 * n].view(dtype).view(shape)

    return dict(final_T=view(0, W * H, torch.float32, (H, W)),
                n_contrib=view(1, W * H, torch.int32, (H, W)),
                ranges=view(2, 2 * T, torch.int32, (T, 2)))


def geom_views(geomBuffer, P):
    """Views of the preprocess outputs inside a geometry buffer:
    dict(means2D [P,2], depths [P], conic_opacity [P,4], rgb [P,3], tiles [P])."""
    P = int(P)
    offs = (ctypes.c_size_t * 5)()
    _lib.load().fgs_geom_offsets(P, offs)

    def view(i, nbytes_per, dtype, shape):
        o = int(offs[i])
        return geomBuffer[o:o + nbytes_per * P].view(dtype).view(shape)

    return dict(means2D=view(0, 8, torch.float32, (P, 2)), depths=view(1, 4, torch.float32, (P,)),
                conic_opacity=view(2, 16, torch.float32, (P, 4)),
                rgb=view(3, 12, torch.float32, (P, 3)), tiles=view(4, 4, torch.int32, (P,)))
