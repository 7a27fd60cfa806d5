"""GPU parity of the gfx950 kernels against the CPU oracle (through the C ABI).

Bars:
  * integer results (radius, tiles touched, depth keys, sorted kv list, tile
    ranges, LoD mask, frustum flags) and the FP32 preprocess outputs that
    feed them: bit-exact against the oracle's FP32 emulation;
  * image: |GPU - oracle FP32| <= 2e-5 except on at most 10 + 1e-5*Npix
    pixels (blend decisions flipped by the hardware exp, each bounded by
    0.01), and |GPU - oracle FP64| <= 1e-4 on every pixel whose blend
    decisions (alpha < 1/255 skips, the T < 1e-4 stop) agree; pixels where
    FP32 and FP64 take a different decision must stay below 0.1 %;
  * gradients: relative L2 error vs the FP64 oracle <= 1e-4 per tensor and
    max-abs <= 2e-3 * max |oracle|.
"""
import numpy as np
import pytest
import torch

from filtergs_amd import ext, workloads
from filtergs_amd.camera import make_camera
from tests._scenes import orc_settings, small_scene

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _t(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _cam_tensors(cam):
    return dict(viewmatrix=_t(cam.world_view_transform), projmatrix=_t(cam.full_proj_transform),
                campos=_t(cam.center))


def gpu_forward(g, cam, variant=0, path="shs", sh_degree=3, bg=(0.0, 0.0, 0.0), stats=False,
                use_filter=True, grad_mode=0, scale_modifier=1.0, **ogfs):
    P = g["means3D"].shape[0]
    ct = _cam_tensors(cam)
    sh = _t(g["shs"]) if path == "shs" else torch.Tensor([])
    colors = _t(g["colors"]) if path == "colors" else torch.Tensor([])
    if path == "cov3D":
        sh = _t(g["shs"])
    cov = _t(g["cov3D"]) if "cov3D" in g and path == "cov3D" else torch.Tensor([])
    scales = torch.Tensor([]) if path == "cov3D" else _t(g["scales"])
    rots = torch.Tensor([]) if path == "cov3D" else _t(g["rotations"])
    out = ext.rasterize_gaussians(
        _t(bg), _t(g["means3D"]), colors, _t(g["opacities"].reshape(P, 1)), scales, rots,
        scale_modifier, cov, ct["viewmatrix"], ct["projmatrix"], cam.tanfovx, cam.tanfovy, cam.height, cam.width, sh,
        sh_degree, ct["campos"], False, use_filter, False, ogfs_enable_stats=stats,
        variant=variant, grad_mode=grad_mode, **ogfs)
    return out


def _np(t):
    return t.detach().cpu().numpy()


def _check_forward(orc, g, cam, variant, path="shs", bg=(0.1, 0.2, 0.3), f64=True, **kw):
    out = gpu_forward(g, cam, variant=variant, path=path, bg=bg, **kw)
    K, color, point_id, wpix, pw, radii, geom, binning, img, _stats = out
    s = orc_settings(orc, cam, sh_degree=3, variant=variant, bg=bg, **kw)
    args = dict(scales=None if path == "cov3D" else g["scales"],
                rotations=None if path == "cov3D" else g["rotations"],
                shs=g["shs"] if path in ("shs", "cov3D") else None,
                colors=g["colors"] if path == "colors" else None,
                cov3D=g.get("cov3D") if path == "cov3D" else None)
    o32 = orc.render(s, g["means3D"], g["opacities"], mode="f32", nthreads=8,
                     decision_hash=True, **args)
    P = g["means3D"].shape[0]
    pre = o32["pre32"]
    # integer stage and its FP32 inputs: bit-exact
    assert np.array_equal(_np(radii), pre["radii"])
    gv = ext.geom_views(geom, P)
    vis = pre["radii"] > 0
    assert np.array_equal(_np(gv["tiles"]).astype(np.uint32), pre["tiles_touched"])
    assert np.array_equal(_np(gv["means2D"])[vis].view(np.uint32), pre["means2D"][vis].view(np.uint32))
    assert np.array_equal(_np(gv["depths"])[vis].view(np.uint32), pre["depths"][vis].view(np.uint32))
    assert np.array_equal(_np(gv["conic_opacity"])[vis].view(np.uint32),
                          pre["conic_opacity"][vis].view(np.uint32))
    assert K == len(o32["vals"])
    keys, vals = ext.read_sorted(geom, P, K, binning)
    assert np.array_equal(_np(keys).view(np.uint64), o32["keys"])
    assert np.array_equal(_np(vals).view(np.uint32), o32["vals"])
    ranges = ext.image_ranges(img, cam.width, cam.height)
    assert np.array_equal(_np(ranges).view(np.uint32), o32["ranges"])
    # colours (contraction-free FP32 in both): tight
    if path != "colors":
        np.testing.assert_allclose(_np(gv["rgb"])[vis], pre["rgb"][vis], atol=2e-6)
    # image vs FP32 emulation: the GPU blend uses the hardware exp, so a pixel
    # can take a different alpha < 1/255 / T < 1e-4 decision than the oracle
    # when a value sits within an ulp of a threshold; such pixels are rare and
    # their difference is bounded by one blend step.
    gimg = _np(color)
    d32 = np.abs(gimg - o32["color"]).max(0)
    agree = d32 <= 2e-5
    assert (~agree).sum() <= 10 + 1e-5 * agree.size, int((~agree).sum())
    assert d32.max() <= 0.01, float(d32.max())
    side = _check_side_outputs(o32, out, img, cam, agree)
    res = dict(out=out, o32=o32, gpu_f32_flips=int((~agree).sum()), side=side)
    if f64:
        o64 = orc.render(s, g["means3D"].astype(np.float64), g["opacities"].astype(np.float64),
                         mode="f64", nthreads=8, decision_hash=True,
                         **{k: (v.astype(np.float64) if v is not None else None) for k, v in args.items()})
        same = (o64["decision_hash"] == o32["decision_hash"]) & agree
        flips = int((~same).sum())
        d = np.abs(gimg - o64["color"]).max(0)
        assert d[same].max() <= 1e-4, d[same].max()
        assert flips <= 1e-3 * same.size, flips
        res["o64"] = o64
        res["maxabs_f64"] = float(d[same].max())
        res["flips"] = flips
    return res


def _pixel_weights(o32, W, px, py):
    """T*alpha of every Gaussian that blends into pixel (px, py), walking its
    tile's sorted list in FP64 with the blend rules of forward.cu:406-567
    (skip power > 0 and alpha < 1/255, stop before T drops below 1e-4)."""
    gx = (W + 15) // 16
    t = (py // 16) * gx + px // 16
    a, e = (int(x) for x in o32["ranges"][t])
    m2 = o32["pre32"]["means2D"].astype(np.float64)
    co = o32["pre32"]["conic_opacity"].astype(np.float64)
    T, w = 1.0, {}
    for k in range(a, e):
        g = int(o32["vals"][k])
        dx, dy = m2[g, 0] - px, m2[g, 1] - py
        power = -0.5 * (co[g, 0] * dx * dx + co[g, 2] * dy * dy) - co[g, 1] * dx * dy
        if power > 0:
            continue
        alpha = min(0.99, co[g, 3] * np.exp(power))
        if alpha < 1 / 255:
            continue
        if T * (1 - alpha) < 1e-4:
            break
        w[g] = T * alpha
        T *= 1 - alpha
    return w


def _check_side_outputs(o32, out, img, cam, agree):
    """The blend's other outputs against the oracle's FP32 emulation, on the
    pixels whose colour agrees (so their alpha / T decisions agree):
    n_contrib bit-exact; final_T and weight_pixel (max T*alpha, strict '>'
    keeps the first maximum in depth order, forward.cu:517-525) within FP32
    rounding of the hardware exp; point_id equal, or -- where two Gaussians
    tie within rounding -- the GPU's choice carries the same weight;
    point_weight (the per-Gaussian maximum of T*alpha over all pixels,
    forward.cu:565) within rounding except for Gaussians seen by a pixel whose
    decisions flipped."""
    K, color, point_id, wpix, pw, radii, geom, binning, imgb, _ = out
    W, H = cam.width, cam.height
    iv = ext.image_views(imgb, W, H)
    nc, fT = _np(iv["n_contrib"]).astype(np.uint32), _np(iv["final_T"])
    pid, wp, pwg = _np(point_id), _np(wpix), _np(pw)
    # n_contrib: exact wherever the decisions agree
    bad_nc = agree & (nc != o32["n_contrib"])
    assert bad_nc.sum() <= 2 + 1e-5 * agree.size, int(bad_nc.sum())
    # final transmittance: a product of up to n_contrib factors (1 - alpha);
    # the hardware exp2 of a power with ~1e-6 absolute rounding moves
    # (1 - alpha) by up to alpha / (1 - alpha) * 1e-6 relative.  A pixel whose
    # colour agrees can still take the other side of an alpha >= 1/255 test
    # where its T is already small (that step moves the colour by less than
    # 2e-5): T then differs by one factor (1 - alpha), alpha within rounding
    # of 1/255 (two such steps at most).
    oT = o32["final_T"]
    relT = np.abs(fT - oT) / np.maximum(oT, 1e-30)
    offT = agree & ~bad_nc & (relT > 1e-4)
    assert offT.sum() <= 10 + 2e-4 * agree.size, int(offT.sum())
    assert (relT[offT] <= 2.02 / 255).all(), float(relT[offT].max())
    ok = agree & ~bad_nc & ~offT
    # pixels nothing blends into: id -1, weight 0, T = 1
    none = ok & (o32["point_id"] == -1)
    assert (pid[none] == -1).all() and (wp[none] == 0).all() and (fT[none] == 1.0).all()
    # maximum weight per pixel
    owp = o32["weight_pixel"]
    assert (np.abs(wp[ok] - owp[ok]) <= 1e-4 * owp[ok] + 1e-7).all()
    # point id: equal, or a tie within rounding (checked by walking the list)
    diff = np.argwhere(ok & (pid != o32["point_id"]))
    assert len(diff) <= 20 + 1e-4 * agree.size, len(diff)
    for py, px in diff[:40]:
        w = _pixel_weights(o32, W, int(px), int(py))
        g_gpu, g_orc = int(pid[py, px]), int(o32["point_id"][py, px])
        assert g_gpu in w and g_orc in w, (px, py, g_gpu, g_orc)
        assert abs(w[g_gpu] - w[g_orc]) <= 1e-4 * w[g_orc], (px, py, w[g_gpu], w[g_orc])
    # point_weight: within rounding except for Gaussians blended into a pixel
    # whose decisions differ (at most n_contrib of them per such pixel)
    opw = o32["point_weight"]
    close = np.abs(pwg - opw) <= 1e-4 * opw + 1e-7
    touched = int(o32["n_contrib"][~ok].astype(np.int64).sum() + nc[~ok].astype(np.int64).sum())
    assert (~close).sum() <= touched + 2, (int((~close).sum()), touched)
    assert np.abs(pwg - opw).max() <= 0.01
    assert (pwg >= 0).all()
    return dict(n_contrib_mismatch=int(bad_nc.sum()), final_T_steps=int(offT.sum()),
                point_id_ties=len(diff), point_weight_off=int((~close).sum()),
                final_T_maxrel=float(relT[ok].max()) if ok.any() else 0.0)


@pytest.mark.parametrize("variant", [0, 1])
def test_forward_sh_small(orc, variant):
    g, cam = small_scene(n=6000, W=320, H=240, seed=101 + variant, dist=7.0)
    _check_forward(orc, g, cam, variant)


def test_forward_c0_scene(orc):
    """Configuration 0 (100k Gaussians, 800x800, SH3, dilate)."""
    g, cam = workloads.config("c0")
    _check_forward(orc, g, cam, 1, bg=(0.0, 0.0, 0.0))


def test_forward_colors_and_cov3d_paths(orc):
    g, cam = small_scene(n=4000, W=256, H=200, seed=111, dist=7.0)
    rng = np.random.default_rng(5)
    g["colors"] = rng.uniform(0, 1, size=(4000, 3)).astype(np.float32)
    _check_forward(orc, g, cam, 0, path="colors")
    s = orc_settings(orc, cam, sh_degree=3, variant=0)
    pre = orc.preprocess(s, g["means3D"], g["opacities"], g["scales"], g["rotations"], g["shs"])
    g["cov3D"] = pre["cov3D"]
    _check_forward(orc, g, cam, 0, path="cov3D")


def test_forward_ogfs_settings_and_no_filter(orc):
    g, cam = small_scene(n=4000, W=256, H=192, seed=121, dist=6.0)
    _check_forward(orc, g, cam, 0, ogfs_epsilon_scale=1.17, ogfs_epsilon_max=0.2,
                   ogfs_energy_floor=0.01, ogfs_radius_ratio_min=0.1)
    _check_forward(orc, g, cam, 0, use_filter=False)


def test_odd_sizes_and_depth_ties(orc):
    """Ragged image (partial tiles), identical depths (stable order by index)."""
    g, cam = small_scene(n=3000, W=203, H=117, seed=131, dist=6.0)
    g["means3D"][100:400, 1] = g["means3D"][100, 1]   # same view depth for a slab
    _check_forward(orc, g, cam, 0)


def test_empty_and_all_culled(orc):
    cam = make_camera(64, 48, center=(0, -5, 0))
    g = workloads.gaussians(0, 1.0, 3, 0)
    out = gpu_forward(g, cam)
    assert out[0] == 0 and out[1].shape == (3, 48, 64)
    g = workloads.gaussians(500, 1.0, 3, 1)
    g["means3D"][:, 1] = -10.0        # all behind the camera
    out = gpu_forward(g, cam, bg=(0.25, 0.5, 0.75))
    assert out[0] == 0
    assert (out[5] == 0).all()
    img = _np(out[1])
    assert np.allclose(img, np.array([0.25, 0.5, 0.75])[:, None, None])


def _grads_gpu(out, g, cam, dL, variant, grad_mode, path="shs", bg=(0.1, 0.2, 0.3),
               scale_modifier=1.0):
    K, color, point_id, wpix, pw, radii, geom, binning, img, _ = out
    ct = _cam_tensors(cam)
    P = g["means3D"].shape[0]
    sh = _t(g["shs"]) if path == "shs" else torch.Tensor([])
    colors = _t(g["colors"]) if path == "colors" else torch.Tensor([])
    return ext.rasterize_gaussians_backward(
        _t(bg), _t(g["means3D"]), radii, colors, _t(g["scales"]), _t(g["rotations"]),
        scale_modifier, torch.Tensor([]), ct["viewmatrix"], ct["projmatrix"], cam.tanfovx, cam.tanfovy, _t(dL), sh,
        3, ct["campos"], geom, K, binning, img, False, variant=variant, use_filter=True,
        grad_mode=grad_mode, opacities=_t(g["opacities"].reshape(P)))


@pytest.mark.parametrize("layout", [0, 1])
@pytest.mark.parametrize("variant,grad_mode", [(0, 0), (0, 1), (1, 0)])
def test_backward_vs_oracle_f64(orc, variant, grad_mode, layout):
    """Both backward tile layouts (four waves per tile / one wave per tile)."""
    from filtergs_amd import _lib
    g, cam = small_scene(n=3000, W=256, H=192, seed=141 + variant, dist=6.0)
    before = _lib.get_backward_layout()
    try:
        _lib.set_backward_layout(layout)
        _check_backward(orc, g, cam, variant, grad_mode)
    finally:
        _lib.set_backward_layout(before)


@pytest.mark.parametrize("layout", [0, 1])
def test_backward_large_gaussians_vs_oracle(orc, layout):
    """Gaussians of more than 16 tiles (their per-tile sums gathered by the
    whole wave) and of more than 128 tiles (added with float atomics) next to
    small ones, in both backward tile layouts."""
    from filtergs_amd import _lib
    g, cam = small_scene(n=3000, W=256, H=192, seed=151, dist=6.0)
    rng = np.random.default_rng(3)
    mid = rng.choice(3000, 60, replace=False)
    g["scales"][mid] = rng.uniform(0.15, 0.35, (60, 3)).astype(np.float32)
    g["scales"][mid[:3]] = 1.2
    g["opacities"][mid] = np.float32(0.3)
    before = _lib.get_backward_layout()
    try:
        _lib.set_backward_layout(layout)
        out = _check_backward(orc, g, cam, 0, 0)
    finally:
        _lib.set_backward_layout(before)
    tiles = _np(ext.geom_views(out[6], 3000)["tiles"]).astype(np.int64)
    assert ((tiles > 16) & (tiles <= 128)).sum() >= 10, np.sort(tiles)[-70:]
    assert (tiles > 128).sum() >= 1


GRAD_NAMES = ["dL_dmean2D", "dL_dcolors", "dL_dopacity", "dL_dmeans3D", "dL_dcov3D", "dL_dsh",
              "dL_dscales", "dL_drot"]


def _check_backward(orc, g, cam, variant, grad_mode, scale_modifier=1.0, dL=None, bg=(0.1, 0.2, 0.3)):
    out = gpu_forward(g, cam, variant=variant, bg=bg, grad_mode=grad_mode,
                      scale_modifier=scale_modifier)
    if dL is None:
        rng = np.random.default_rng(7)
        dL = rng.normal(size=(3, cam.height, cam.width)).astype(np.float32)
    gg = _grads_gpu(out, g, cam, dL, variant, grad_mode, bg=bg, scale_modifier=scale_modifier)
    _compare_grads(orc, g, cam, dL, dict(zip(GRAD_NAMES, gg)), variant, grad_mode, bg,
                   scale_modifier)
    return out


def _compare_grads(orc, g, cam, dL, got, variant, grad_mode, bg, scale_modifier=1.0, path="shs"):
    """Each GPU gradient in ``got`` (oracle names) against the FP64 oracle of
    the same scene and dL/dpix."""
    s = orc_settings(orc, cam, sh_degree=3, variant=variant, bg=bg, grad_mode=grad_mode,
                     scale_modifier=scale_modifier)

    def oracle_grads(mode):
        dt = np.float64 if mode == "f64" else np.float32
        d = {k: g[k].astype(dt) for k in g}
        col = dict(colors=d["colors"]) if path == "colors" else dict(shs=d["shs"])
        f = orc.render(s, d["means3D"], d["opacities"], d["scales"], d["rotations"],
                       mode=mode, nthreads=8, **col)
        ob = orc.backward(s, f, dL.astype(dt), d["means3D"], d["opacities"], d["scales"],
                          d["rotations"], mode=mode, **col)
        return dict(dL_dmean2D=ob["dL_dmean2D"], dL_dcolors=ob["dL_dcolors"],
                    dL_dopacity=ob["dL_dopacity"].reshape(-1, 1), dL_dmeans3D=ob["dL_dmeans3D"],
                    dL_dcov3D=ob["dL_dcov3D"], dL_dsh=ob["dL_dsh"], dL_dscales=ob["dL_dscales"],
                    dL_drot=ob["dL_drot"])

    ref = oracle_grads("f64")
    ref32 = oracle_grads("f32")

    def rel(a, r):
        return np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30)

    # FP32 atomics vs FP64: relative L2 error per tensor within 1e-4, or -- for
    # the tensors whose chain has cancellations (quaternion / scale gradients
    # of sub-pixel Gaussians) -- no worse than 3x the error of the sequential
    # FP32 oracle itself.  The max-abs bound leaves room for the rare pixel
    # whose blend decision differs between the hardware exp and FP64 (one
    # step: alpha*T*dL <= ~1/255 * |dL|).
    for name, t in got.items():
        a = _np(t).astype(np.float64)
        r = ref[name].reshape(a.shape)
        r32 = ref32[name].reshape(a.shape).astype(np.float64)
        scale = np.abs(r).max()
        e_gpu, e_f32 = rel(a, r), rel(r32, r)
        err = np.abs(a - r).max()
        assert e_gpu <= max(1e-4, 3 * e_f32 + 1e-5), (name, e_gpu, e_f32)
        assert err <= 2e-3 * scale + 1e-9, (name, err, scale)


@pytest.mark.parametrize("variant,grad_mode", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_c2a_backward_vs_oracle_and_fd(orc, variant, grad_mode):
    """Configuration 2a (4096 Gaussians in [-1,1]^3 at 128x128): every GPU
    gradient against the FP64 oracle; in exact mode also the GPU gradient of
    a sample of means, scales, rotations, opacities and SH coefficients
    against FP64 central differences (eps 1e-6) of the oracle forward, the
    tile structure held fixed and samples whose perturbation flips a blend
    decision skipped."""
    g, cam = workloads.config("c2a", seed=71 + variant)
    _check_backward(orc, g, cam, variant, grad_mode)
    if grad_mode != 1:
        return
    from tests.test_oracle_grad import _Problem
    prob = _Problem(orc, variant, "shs", seed=71 + variant, scene=(g, cam))
    bg = (0.2, 0.5, 0.1)
    out = gpu_forward(g, cam, variant=variant, bg=bg, grad_mode=1)
    gg = _grads_gpu(out, g, cam, prob.dL.astype(np.float32), variant, 1, bg=bg)
    gpu = dict(means3D=_np(gg[3]), opacities=_np(gg[2]).reshape(-1), scales=_np(gg[6]),
               rotations=_np(gg[7]), shs=_np(gg[5]))
    rng = np.random.default_rng(73)
    vis = np.flatnonzero(prob.radii > 0)
    assert vis.size > 3000
    checked = 0
    for name, a in gpu.items():
        per = a.reshape(a.shape[0], -1).shape[1]
        flat = a.reshape(-1).astype(np.float64)
        scale = np.abs(flat).max()
        for gi in rng.choice(vis, 8, replace=False):
            for comp in rng.choice(per, min(per, 4), replace=False):
                idx = int(gi) * per + int(comp)
                fd = prob.fd(name, idx)
                if fd is None:
                    continue
                assert abs(flat[idx] - fd) <= 1e-3 * abs(fd) + 1e-4 * scale, (name, gi, comp, flat[idx], fd)
                checked += 1
    assert checked >= 100


@pytest.mark.parametrize("stats", [False, True])
def test_blend_walks_agree(stats):
    """The forward's wave-uniform walk (DPP maximum), its divergent walk
    (per-lane LDS maxima) and the adaptive choice between them render the
    same bits: image, point ids, pixel weights, point weights (with the
    statistics on too; their per-pair float sums meet in LDS atomics, whose
    order is not fixed, so the statistics are checked against the oracle in
    test_stats_vs_oracle instead).  Two scenes: large overlapping Gaussians
    (most lanes blend at each entry) and a far view of many small ones (few
    do), both with tiles of several staging rounds."""
    from filtergs_amd import _lib
    scenes = [small_scene(n=20000, W=320, H=240, seed=5, dist=4.0),
              small_scene(n=60000, W=320, H=240, seed=6, dist=14.0)]
    try:
        _lib.set_forward_layout(0)   # the walk policy belongs to the four-wave layout
        for g, cam in scenes:
            outs = []
            for mode in (0, 1, 2):
                _lib.set_blend_walk(mode)
                outs.append(gpu_forward(g, cam, stats=stats))
            _lib.set_forward_layout(1)
            if not stats:   # the one-wave layout: the same bits again
                outs.append(gpu_forward(g, cam, stats=stats))
            _lib.set_forward_layout(0)
            ranges = _np(ext.image_ranges(outs[0][8], cam.width, cam.height)).astype(np.int64)
            assert (ranges[:, 1] - ranges[:, 0]).max() > 256   # several rounds in some tile
            for o in outs[1:]:
                for k in (1, 2, 3, 4, 5):
                    assert torch.equal(o[k], outs[0][k]), k
    finally:
        _lib.set_blend_walk(0)
        _lib.set_forward_layout(2)


@pytest.mark.parametrize("grad_mode,path", [("reference", "shs"), ("exact", "shs"),
                                            ("reference", "colors")])
def test_autograd_through_plugin_surface(orc, grad_mode, path):
    """GaussianRasterizer forward / backward through torch.autograd: every
    leaf's .grad against the FP64 oracle for loss = sum(image * dL), so a
    gradient routed to the wrong input (means3D vs means2D, SH vs colours)
    fails."""
    from filtergs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer
    g, cam = small_scene(n=2000, W=160, H=128, seed=151, dist=6.0)
    bg = (0.1, 0.2, 0.3)
    rng = np.random.default_rng(152)
    g["colors"] = rng.uniform(0, 1, size=(2000, 3)).astype(np.float32)
    ct = _cam_tensors(cam)
    rs = GaussianRasterizationSettings(
        image_height=cam.height, image_width=cam.width, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy,
        bg=_t(bg), scale_modifier=1.0, viewmatrix=ct["viewmatrix"],
        projmatrix=ct["projmatrix"], sh_degree=3, campos=ct["campos"], prefiltered=False,
        debug=False, grad_mode=grad_mode)
    r = GaussianRasterizer(rs)
    keys = ("means3D", "scales", "rotations") + (("shs",) if path == "shs" else ("colors",))
    params = {k: _t(g[k]).requires_grad_(True) for k in keys}
    opac = _t(g["opacities"].reshape(-1, 1)).requires_grad_(True)
    means2D = torch.zeros_like(params["means3D"], requires_grad=True)
    color, radii, pid, wpix, pw = r(means3D=params["means3D"], means2D=means2D, opacities=opac,
                                    shs=params.get("shs"), colors_precomp=params.get("colors"),
                                    scales=params["scales"], rotations=params["rotations"])
    assert r.latest_stats is not None and r.latest_stats["num_gaussians"] == 2000
    assert "tile_ratio" in r.latest_stats and len(r.latest_stats["tile_ratio"]) == r.latest_stats["num_tiles"]
    dL = rng.normal(size=(3, cam.height, cam.width)).astype(np.float32)
    (color * _t(dL)).sum().backward()
    got = dict(dL_dmean2D=means2D.grad, dL_dopacity=opac.grad, dL_dmeans3D=params["means3D"].grad,
               dL_dscales=params["scales"].grad, dL_drot=params["rotations"].grad)
    if path == "shs":
        got["dL_dsh"] = params["shs"].grad
    else:
        got["dL_dcolors"] = params["colors"].grad
    shapes = dict(dL_dmean2D=(2000, 3), dL_dopacity=(2000, 1), dL_dmeans3D=(2000, 3),
                  dL_dscales=(2000, 3), dL_drot=(2000, 4), dL_dsh=(2000, 16, 3), dL_dcolors=(2000, 3))
    for k, v in got.items():
        assert v is not None and tuple(v.shape) == shapes[k], (k, None if v is None else v.shape)
    _compare_grads(orc, g, cam, dL, got, 0, 1 if grad_mode == "exact" else 0, bg, path=path)


@pytest.mark.parametrize("grad_mode", ["reference", "exact"])
def test_dilate_plugin_surface_vs_oracle(orc, grad_mode):
    """The dilate plugin (antialias/diff_gaussian_rasterization/__init__.py)
    forward and backward through torch.autograd against oracle variant 1
    (+0.3 dilation, 3-sigma radius): radii bit-exact, image and point weights
    against the FP32 emulation, every gradient against FP64; and its _C
    called directly with the 19-argument signature returns the same 8
    values."""
    from filtergs_amd import dilate, dilate_ext
    g, cam = small_scene(n=3000, W=192, H=144, seed=221, dist=6.0)
    bg = (0.1, 0.2, 0.3)
    ct = _cam_tensors(cam)
    rs = dilate.GaussianRasterizationSettings(
        image_height=cam.height, image_width=cam.width, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy,
        bg=_t(bg), scale_modifier=1.0, viewmatrix=ct["viewmatrix"], projmatrix=ct["projmatrix"],
        sh_degree=3, campos=ct["campos"], prefiltered=False, debug=False, grad_mode=grad_mode)
    params = {k: _t(g[k]).requires_grad_(True) for k in ("means3D", "scales", "rotations", "shs")}
    opac = _t(g["opacities"].reshape(-1, 1)).requires_grad_(True)
    means2D = torch.zeros_like(params["means3D"], requires_grad=True)
    color, radii, pid, pw = dilate.GaussianRasterizer(rs)(
        means3D=params["means3D"], means2D=means2D, opacities=opac, shs=params["shs"],
        scales=params["scales"], rotations=params["rotations"])
    s = orc_settings(orc, cam, sh_degree=3, variant=1, bg=bg)
    o32 = orc.render(s, g["means3D"], g["opacities"], g["scales"], g["rotations"], g["shs"],
                     mode="f32", nthreads=8)
    assert np.array_equal(_np(radii), o32["pre32"]["radii"])
    d32 = np.abs(_np(color) - o32["color"]).max(0)
    assert (d32 > 2e-5).sum() <= 10 + 1e-5 * d32.size and d32.max() <= 0.01
    opw = o32["point_weight"]
    off = np.abs(_np(pw) - opw) > 1e-4 * opw + 1e-7
    assert off.sum() <= 4 * int((d32 > 2e-5).sum()) + 2
    agree = d32 <= 2e-5
    assert (_np(pid)[agree] == o32["point_id"][agree]).mean() > 0.999
    rng = np.random.default_rng(222)
    dL = rng.normal(size=(3, cam.height, cam.width)).astype(np.float32)
    (color * _t(dL)).sum().backward()
    got = dict(dL_dmean2D=means2D.grad, dL_dopacity=opac.grad, dL_dmeans3D=params["means3D"].grad,
               dL_dscales=params["scales"].grad, dL_drot=params["rotations"].grad,
               dL_dsh=params["shs"].grad)
    _compare_grads(orc, g, cam, dL, got, 1, 1 if grad_mode == "exact" else 0, bg)
    # the _C entry with graphdeco's 19 positional arguments, 8 results
    with torch.no_grad():
        out = dilate_ext.rasterize_gaussians(
            rs.bg, params["means3D"], torch.Tensor([]), opac, params["scales"], params["rotations"],
            1.0, torch.Tensor([]), rs.viewmatrix, rs.projmatrix, rs.tanfovx, rs.tanfovy,
            rs.image_height, rs.image_width, params["shs"], 3, rs.campos, False, False)
    assert len(out) == 8
    num_rendered, c2, pid2, pw2, radii2 = out[:5]
    assert num_rendered == len(o32["vals"])
    assert torch.equal(c2, color) and torch.equal(pid2, pid) and torch.equal(pw2, pw)
    assert torch.equal(radii2, radii)
    gr = dilate_ext.rasterize_gaussians_backward(
        rs.bg, params["means3D"], radii2, torch.Tensor([]), params["scales"], params["rotations"],
        1.0, torch.Tensor([]), rs.viewmatrix, rs.projmatrix, rs.tanfovx, rs.tanfovy, _t(dL),
        params["shs"], 3, rs.campos, out[5], num_rendered, out[6], out[7], False,
        grad_mode=1 if grad_mode == "exact" else 0, opacities=opac.detach())
    assert len(gr) == 8
    d = (gr[3].double() - params["means3D"].grad.double()).norm() / params["means3D"].grad.double().norm()
    assert float(d) <= 1e-5


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("grad_mode", [0, 1])
@pytest.mark.parametrize("scale_modifier", [0.5, 1.7])
def test_backward_clamp_and_scale_modifier(orc, variant, grad_mode, scale_modifier):
    """Half of the Gaussians with opacity in [0.99, 1) and three times the
    size, so o*G passes 0.99 at the pixels next to their centres: the 0.99
    alpha clamp is hit (the reference backward passes the gradient through
    it, the exact mode stops it, backward.cu:499,538).  scale_modifier != 1:
    the reference's dL/dscale is taken w.r.t. mod*scale and not multiplied by
    the modifier (backward.cu:295,323-325); the exact mode is.  Both grad
    modes, both variants, against the FP64 oracle."""
    n = 2500
    g, cam = small_scene(n=n, W=224, H=160, seed=211 + variant, dist=6.0)
    rng = np.random.default_rng(212)
    hi = rng.choice(n, n // 2, replace=False)
    g["opacities"][hi, 0] = rng.uniform(0.99, 0.9999, hi.size).astype(np.float32)
    g["scales"][hi] *= np.float32(3.0)
    _check_backward(orc, g, cam, variant, grad_mode, scale_modifier)
    # the clamp is reached: o*G >= 0.99 at the pixel nearest to a centre
    s = orc_settings(orc, cam, sh_degree=3, variant=variant, scale_modifier=scale_modifier)
    pre = orc.preprocess(s, g["means3D"], g["opacities"], g["scales"], g["rotations"], g["shs"])
    vis = np.flatnonzero(pre["radii"] > 0)
    m = pre["means2D"][vis].astype(np.float64)
    pix = np.rint(m)
    d = m - pix
    co = pre["conic_opacity"][vis].astype(np.float64)
    G = np.exp(-0.5 * (co[:, 0] * d[:, 0] ** 2 + co[:, 2] * d[:, 1] ** 2) - co[:, 1] * d[:, 0] * d[:, 1])
    inside = (pix[:, 0] >= 0) & (pix[:, 0] < cam.width) & (pix[:, 1] >= 0) & (pix[:, 1] < cam.height)
    assert ((co[:, 3] * G >= 0.99) & inside).sum() >= 20


@pytest.mark.parametrize("case", ["sparse", "saturated"])
def test_stats_vs_oracle(orc, case, monkeypatch):
    """RenderStats computed inside the blend against the oracle's per-kv-pair
    sums and counts, binned and summed as rasterizer_impl.cu:479-556 does.
    "saturated": opaque Gaussians, so tiles stop before their last batch (the
    pairs never staged count as unused), and a KPC_USED_THRESH override."""
    thr = 1e-4
    if case == "sparse":
        g, cam = small_scene(n=3000, W=256, H=192, seed=161, dist=6.0)
    else:
        g, cam = small_scene(n=8000, W=256, H=192, seed=7, dist=4.0)
        g["opacities"] = np.full_like(g["opacities"], 0.97)
        thr = 0.05
        monkeypatch.setenv("KPC_USED_THRESH", "0.05")
    out = gpu_forward(g, cam, variant=0, stats=True)
    from filtergs_amd.rasterizer import parse_stats
    st = parse_stats(out[9])
    s = orc_settings(orc, cam, sh_degree=3, variant=0)
    f = orc.render(s, g["means3D"], g["opacities"], g["scales"], g["rotations"], g["shs"],
                   mode="f32", stats=True)
    kv, cnt, rg = f["kv_alpha_sum"].astype(np.float64), f["kv_alpha_count"], f["ranges"]
    T = rg.shape[0]
    used = cnt > 0
    if case == "saturated":
        # some tile ends with pairs that no pixel reached
        assert any(e > a and cnt[e - 1] == 0 for a, e in rg)
    edges = np.array([0.01, 0.05, 0.10, 0.15, 0.20, 0.25, 0.30, 0.35, 0.40, 0.45, 0.50], np.float32)
    b = np.searchsorted(edges, f["kv_alpha_sum"], side="left")
    inb = used & (b < 11)
    assert st["num_kv_pairs"] == len(kv)
    assert st["unused_kv_pairs"] == int((~used).sum())
    assert abs(st["used_kv_pairs"] - int(inb.sum())) <= 2
    hist = np.bincount(b[inb], minlength=11)
    assert np.abs(np.array(st["kpc_hist"]) - hist).sum() <= 4
    tile_avg = np.array([kv[a:e].mean() if e > a else 0.0 for a, e in rg])
    assert abs(st["bar_g"] - tile_avg.sum() / T) <= 1e-5
    np.testing.assert_allclose(st["tile_total_kv"], [e - a for a, e in rg])
    ratio = np.array([((cnt[a:e] > 0) & (f["kv_alpha_sum"][a:e] >= np.float32(thr))).mean()
                      if e > a else 0.0 for a, e in rg])
    # a pair sum within an ulp of the threshold may land on either side
    dr = np.abs(np.array(st["tile_ratio"]) - ratio)
    assert dr.max() <= 0.1 and dr.sum() <= 0.05, (dr.max(), dr.sum())


def test_lod_cut_bitexact(orc):
    from filtergs_amd import lod
    from tests.test_oracle_lod import random_tree
    tree = random_tree(n_roots=200, depth=6, seed=3)
    cam = make_camera(640, 480, center=(0.0, -9.0, 2.0), fovx_deg=70.0)
    fx = cam.width / (2 * cam.tanfovx)
    fy = cam.height / (2 * cam.tanfovy)
    for rmax, thr in ((5.0, 700.0), (1.5, 40.0)):
        mask, r2d, depth = lod.compute_visibility_radius(
            _t(tree["means3D"]), _t(tree["log_scales"]), _t(tree["rotations"]),
            _t(cam.full_proj_transform), _t(cam.world_view_transform),
            _t(tree["node_index"], torch.int32), _t(tree["ancestor_path"], torch.int32), fx, fy,
            cam.tanfovx, cam.tanfovy, rmax, padding=0.1, radius_filter_threshold=thr)
        om, orr, od = orc.lod_cut(tree["means3D"], tree["log_scales"], tree["rotations"],
                                  cam.full_proj_transform, cam.world_view_transform,
                                  tree["node_index"], tree["ancestor_path"], fx, fy, cam.tanfovx,
                                  cam.tanfovy, 0.1, rmax, thr)
        assert np.array_equal(_np(mask), om)
        assert np.array_equal(_np(r2d).view(np.uint32), orr.view(np.uint32))
        assert np.array_equal(_np(depth).view(np.uint32), od.view(np.uint32))
        assert 0 < om.sum() < len(om)


def test_point_utilities(orc):
    from filtergs_amd import lod
    g, cam = small_scene(n=5000, W=320, H=240, seed=171, dist=6.0, half=4.0)
    m, sc, ro = _t(g["means3D"]), _t(g["scales"]), _t(g["rotations"])
    vm, pm = _t(cam.world_view_transform), _t(cam.full_proj_transform)
    vis = ext.mark_visible(m, vm, pm)
    assert np.array_equal(_np(vis), orc.mark_visible(g["means3D"], cam.world_view_transform,
                                                     cam.full_proj_transform))
    ra = ext.compute_radius(m, sc, ro, vm, pm, cam.tanfovx, cam.tanfovy, cam.height, cam.width)
    assert np.array_equal(_np(ra).view(np.uint32),
                          orc.radius_aa(g["means3D"], g["scales"], g["rotations"],
                                        cam.world_view_transform, cam.full_proj_transform,
                                        cam.tanfovx, cam.tanfovy, cam.height, cam.width).view(np.uint32))
    fx, fy = cam.width / (2 * cam.tanfovx), cam.height / (2 * cam.tanfovy)
    rv = lod.compute_radius_module.compute_radius(m, sc, ro, pm, vm, fx, fy, cam.tanfovx, cam.tanfovy)
    assert np.array_equal(_np(rv).view(np.uint32),
                          orc.radius_v2(g["means3D"], g["scales"], g["rotations"],
                                        cam.full_proj_transform, cam.world_view_transform, fx, fy,
                                        cam.tanfovx, cam.tanfovy).view(np.uint32))
    valid, depth, pp = lod.visible_flag_cuda_direct(m, pm, padding=0.05)
    ov, od, op = orc.visible_flag(g["means3D"], cam.full_proj_transform, 0.05)
    assert np.array_equal(_np(valid), ov)
    assert np.array_equal(_np(depth).view(np.uint32), od.view(np.uint32))
    assert np.array_equal(_np(pp).view(np.uint32), op.view(np.uint32))
    vb, db, pb = lod.visible_flag_cuda_batch_direct(torch.stack([m, m * 0.5]), pm, padding=0.05)
    assert np.array_equal(_np(vb[0]), ov)
    idx = torch.tensor([4, 0, 4999, 17], dtype=torch.int64, device=DEV)
    cols = _t(np.random.default_rng(1).uniform(size=(5000, 3)))
    shs = _t(g["shs"][:, 1:, :])
    out = lod.gather_attributes(idx, m, sc, ro, _t(g["opacities"].reshape(-1, 1)), cols, shs)
    for k, src in (("xyz", m), ("scaling", sc), ("rotation", ro), ("colors", cols), ("shs", shs)):
        assert torch.equal(out[k], src[idx]), k
    with pytest.raises(RuntimeError):
        lod.gather_attributes(torch.tensor([5000], device=DEV), m, sc, ro,
                              _t(g["opacities"].reshape(-1, 1)), cols)


def test_c1_scale_keys_bitexact_and_image(orc):
    """Configuration 1 (1M Gaussians in a 40^3 box, 1920x1080, SH3): tile keys,
    sort order and ranges bit-exact; image against FP64."""
    g, cam = workloads.config("c1")
    res = _check_forward(orc, g, cam, 0, bg=(0.0, 0.0, 0.0))
    assert res["out"][0] > 1_000_000


def test_work_shapes_agree():
    """Every blend work shape (Gaussians per LDS batch, forward and backward;
    both tile layouts of the forward and of the backward) renders the same image, n_contrib, max-weight ids and point weights bit
    for bit, and the same gradients up to float summation order.  The scene
    puts several hundred Gaussians on most tiles so every shape runs several
    batches."""
    from filtergs_amd import _lib
    g, cam = small_scene(n=40000, W=200, H=136, seed=161, dist=5.0)
    rng = np.random.default_rng(8)
    dL = rng.normal(size=(3, cam.height, cam.width)).astype(np.float32)
    bg = (0.1, 0.2, 0.3)
    base = None
    try:
        for fb, bb, lay, flay in ((256, 128, 1, 2), (128, 64, 0, 1), (256, 256, 0, 0), (128, 128, 0, 1),
                                  (256, 128, 0, 0), (256, 128, 1, 1)):
            _lib.set_work_shape(fb, bb)
            _lib.set_backward_layout(lay)
            _lib.set_forward_layout(flay)
            assert _lib.get_work_shape() == (fb, bb)
            out = gpu_forward(g, cam, bg=bg)
            gr = _grads_gpu(out, g, cam, dL, 0, 0, bg=bg)
            iv = ext.image_views(out[8], cam.width, cam.height)
            fw = [_np(out[k]) for k in (1, 2, 3, 4)] + [_np(iv[k]) for k in
                                                        ("final_T", "n_contrib", "ranges")]
            res = (fw, [_np(t) for t in gr])
            if base is None:
                base = res
                assert int(out[0]) > 200 * (cam.width // 16) * (cam.height // 16)
                continue
            for k, (a, b) in enumerate(zip(res[0], base[0])):
                assert np.array_equal(a, b), (fb, bb, "forward output", k)
            for k, (a, b) in enumerate(zip(res[1], base[1])):
                # blend terms (mean2D, colors, opacity) differ by atomic
                # summation order only; the chain through cov3D / rotation
                # has cancellations that amplify it
                d = np.linalg.norm(a.astype(np.float64) - b) / max(np.linalg.norm(b), 1e-30)
                assert d <= (1e-5 if k < 3 else 1e-4), (fb, bb, "gradient", k, d)
    finally:
        _lib.set_work_shape(256, 128)
        _lib.set_backward_layout(1)
        _lib.set_forward_layout(2)
    with pytest.raises(Exception):
        _lib.set_work_shape(512, 128)
    with pytest.raises(Exception):
        _lib.set_backward_layout(2)
    with pytest.raises(Exception):
        _lib.set_forward_layout(3)
