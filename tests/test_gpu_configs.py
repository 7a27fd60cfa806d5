"""The benchmark configurations at their full sizes on the GPU:

* 2b -- a training step of 3M Gaussians at 1920x1080 with LoG's loss
  (0.8 L1 + 0.2 (1 - SSIM), LoG/render/renderer.py:260-273) through the
  plugin's autograd; the loss's image gradient on a 256-row band checked
  against the FP64 oracle;
* 3 -- the 16 orbit cameras over 3M Gaussians (the benchmark step): the
  batched path equal to 16 plain per-view calls, and a band of a 3840x2160
  view against the oracle;
* 4 -- the LoD cut of the 22.4M-node urban tree for an aerial view, bit-exact
  against the oracle (ancestor rows and parent links).
"""
import numpy as np
import pytest
import torch

from filtergs_amd import ext, workloads
from tests.test_gpu_parity import _check_backward, _check_forward, _np, _t

pytestmark = pytest.mark.gpu


def test_c2b_training_step(orc):
    from filtergs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer
    g, cam = workloads.config("c2b")
    H, W = cam.height, cam.width
    rs = GaussianRasterizationSettings(
        image_height=H, image_width=W, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy,
        bg=_t([0.0, 0.0, 0.0]), scale_modifier=1.0, viewmatrix=_t(cam.world_view_transform),
        projmatrix=_t(cam.full_proj_transform), sh_degree=3, campos=_t(cam.center),
        prefiltered=False, debug=False)
    r = GaussianRasterizer(rs)
    leaves = {k: _t(g[k]).requires_grad_(True) for k in ("means3D", "scales", "rotations", "shs")}
    leaves["opacities"] = _t(g["opacities"].reshape(-1, 1)).requires_grad_(True)
    means2D = torch.zeros_like(leaves["means3D"], requires_grad=True)
    target = torch.rand((3, H, W), generator=torch.Generator(device="cuda").manual_seed(5),
                        device="cuda")

    def render(p):
        return r(means3D=p["means3D"], means2D=means2D, opacities=p["opacities"], shs=p["shs"],
                 scales=p["scales"], rotations=p["rotations"])[0]

    color = render(leaves)
    color.retain_grad()
    loss = workloads.l1_ssim_loss(color, target)
    loss.backward()
    L0 = float(loss.detach())
    assert np.isfinite(L0) and 0 < L0 < 1
    for k, p in leaves.items():
        assert p.grad is not None and p.grad.shape == p.shape and torch.isfinite(p.grad).all(), k
    assert means2D.grad.abs().sum() > 0
    # first order: a small step along -grad on the SH rows lowers the loss by
    # about eta * |grad|^2
    gsh = leaves["shs"].grad
    gg = float((gsh.double() ** 2).sum())
    eta = 1e-3 * L0 / gg
    with torch.no_grad():
        step = {k: v.detach() for k, v in leaves.items()}
        step["shs"] = step["shs"] - eta * gsh
        L1 = float(workloads.l1_ssim_loss(render(step), target))
    ratio = (L0 - L1) / (eta * gg)
    assert 0.7 < ratio < 1.3, (L0, L1, eta * gg)
    # parity on a band: the loss's image gradient there, fed to the plain
    # path and to the FP64 oracle on the band's camera
    y0, rows = 512, 256
    bcam = workloads.band_camera(cam, y0, rows)
    dL = np.ascontiguousarray(_np(color.grad)[:, y0:y0 + rows, :]).astype(np.float32)
    assert np.abs(dL).max() > 0
    _check_forward(orc, g, bcam, 0, bg=(0.0, 0.0, 0.0))
    _check_backward(orc, g, bcam, 0, 0, dL=dL, bg=(0.0, 0.0, 0.0))


def _plain(g_t, cam, bg, dL):
    means, scales, rots, opac, shs = g_t
    vm, pm, cp = _t(cam.world_view_transform), _t(cam.full_proj_transform), _t(cam.center)
    e = torch.Tensor([])
    out = ext.rasterize_gaussians(bg, means, e, opac, scales, rots, 1.0, e, vm, pm, cam.tanfovx,
                                  cam.tanfovy, cam.height, cam.width, shs, 3, cp, False, True, False,
                                  ogfs_enable_stats=False)
    K, color, pid, wpix, pw, radii, geom, binning, img, _ = out
    gr = ext.rasterize_gaussians_backward(bg, means, radii, e, scales, rots, 1.0, e, vm, pm,
                                          cam.tanfovx, cam.tanfovy, dL, shs, 3, cp, geom, K, binning,
                                          img, False)
    return (K, color, pid, wpix, pw, radii), gr


def test_c3_batch_equals_per_view_full_size(orc):
    from filtergs_amd.batch import BatchRasterizer
    P = 3_000_000
    g = workloads.gaussians(P, 20.0, 3, seed=0)
    cams = workloads.orbit_cameras(16)
    g_t = (_t(g["means3D"]), _t(g["scales"]), _t(g["rotations"]), _t(g["opacities"].reshape(-1, 1)),
           _t(g["shs"]))
    br = BatchRasterizer(cams, P)
    br.bind(g_t[0], g_t[3], g_t[1], g_t[2], shs=g_t[4])
    K = br.forward()
    assert 0 < K <= br.K_cap
    gen = torch.Generator(device="cuda").manual_seed(11)
    dL = torch.randn((3 * int(br.sizes.pixels),), generator=gen, device="cuda")
    grads = br.backward(dL)
    bg = _t([0.0, 0.0, 0.0])
    summed = None
    Ksum = 0
    off = 0
    for v, cam in enumerate(cams):
        n = 3 * cam.height * cam.width
        fwd, gr = _plain(g_t, cam, bg, dL[off:off + n].view(3, cam.height, cam.width))
        off += n
        Ksum += int(fwd[0])
        assert torch.equal(br.radii[v], fwd[5]), v
        assert torch.equal(br.view_color(v), fwd[1]), v
        assert torch.equal(br.view_pixels(br.point_id, v), fwd[2]), v
        assert torch.equal(br.view_pixels(br.weight_pixel, v), fwd[3]), v
        assert torch.equal(br.point_weight[v], fwd[4]), v
        for a, b in zip(br.view_grads(v), (gr[0], gr[1])):
            d = float((a.double() - b.double()).norm() / b.double().norm().clamp_min(1e-30))
            assert d <= 1e-5, (v, d)
        part = [gr[i].double() for i in (2, 3, 5, 6, 7)]
        summed = part if summed is None else [s_ + p_ for s_, p_ in zip(summed, part)]
        del fwd, gr
    assert K == Ksum
    for k, ref in zip(("opacities", "means3D", "shs", "scales", "rotations"), summed):
        got = grads[k].double().reshape(ref.shape)
        d = float((got - ref).norm() / ref.norm().clamp_min(1e-30))
        assert d <= 1e-4, (k, d)
    # a band of the first 3840x2160 view against the oracle
    big = next(c for c in cams if c.width == 3840)
    bcam = workloads.band_camera(big, 1024, 256)
    _check_forward(orc, g, bcam, 0, bg=(0.0, 0.0, 0.0))
    _check_backward(orc, g, bcam, 0, 0, bg=(0.0, 0.0, 0.0))


def test_c4_lod_cut_full_tree(orc):
    from filtergs_amd import lod
    tree = workloads.lod_tree()
    N = int(tree["means3D"].shape[0])
    assert N > 20_000_000
    cam = workloads.aerial_cameras(8)[0]
    fx, fy = cam.width / (2 * cam.tanfovx), cam.height / (2 * cam.tanfovy)
    d = {k: tree[k].cuda() for k in ("means3D", "log_scales", "rotations", "tree_node_index",
                                     "ancestor_path", "index_parent")}
    pm, vm = _t(cam.full_proj_transform), _t(cam.world_view_transform)
    host = {k: tree[k].numpy() for k in ("means3D", "log_scales", "rotations", "tree_node_index",
                                         "ancestor_path")}
    for rmax in (1.5, 5.0):
        mask, r2d, depth = lod.compute_visibility_radius(
            d["means3D"], d["log_scales"], d["rotations"], pm, vm, d["tree_node_index"],
            d["ancestor_path"], fx, fy, cam.tanfovx, cam.tanfovy, rmax, padding=0.1,
            radius_filter_threshold=700.0)
        om, orr, od = orc.lod_cut(host["means3D"], host["log_scales"], host["rotations"],
                                  cam.full_proj_transform, cam.world_view_transform,
                                  host["tree_node_index"], host["ancestor_path"], fx, fy, cam.tanfovx,
                                  cam.tanfovy, 0.1, rmax, 700.0, nthreads=16)
        assert np.array_equal(_np(mask), om), rmax
        assert np.array_equal(_np(r2d).view(np.uint32), orr.view(np.uint32)), rmax
        assert np.array_equal(_np(depth).view(np.uint32), od.view(np.uint32)), rmax
        assert 100_000 < om.sum() < N // 2, (rmax, int(om.sum()))
        mp, rp, dp = lod.compute_visibility_radius_parent(
            d["means3D"], d["log_scales"], d["rotations"], pm, vm, d["tree_node_index"],
            d["index_parent"], fx, fy, cam.tanfovx, cam.tanfovy, rmax, padding=0.1,
            radius_filter_threshold=700.0)
        assert torch.equal(mp, mask) and torch.equal(rp, r2d) and torch.equal(dp, depth), rmax
