"""The batched multi-view path against the plain per-view path (both through
the C ABI): every per-view forward output bit-exact, per-view and summed
gradients equal up to float summation order, graph replays equal to eager
calls, capacity growth."""
import numpy as np
import pytest
import torch

from filtergs_amd import ext
from filtergs_amd.batch import BatchRasterizer
from filtergs_amd.camera import make_camera
from tests._scenes import small_scene

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _t(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _cams():
    return [make_camera(200, 136, center=(0.0, -5.0, 0.5), fovx_deg=60.0),
            make_camera(160, 120, center=(4.0, -3.0, 1.0), fovx_deg=55.0),
            make_camera(96, 80, center=(-3.5, 2.0, -1.0), fovx_deg=70.0),
            make_camera(130, 70, center=(0.5, 5.5, 0.2), fovx_deg=65.0)]


def _single(g, cam, bg, dL, colors=None, grad_mode=0, deg=3):
    P = g["means3D"].shape[0]
    sh = _t(g["shs"]) if colors is None else torch.Tensor([])
    col = _t(colors) if colors is not None else torch.Tensor([])
    vm, pm, cp = _t(cam.world_view_transform), _t(cam.full_proj_transform), _t(cam.center)
    out = ext.rasterize_gaussians(_t(bg), _t(g["means3D"]), col, _t(g["opacities"].reshape(P, 1)),
                                  _t(g["scales"]), _t(g["rotations"]), 1.0, torch.Tensor([]), vm, pm,
                                  cam.tanfovx, cam.tanfovy, cam.height, cam.width, sh, deg, cp, False,
                                  True, False, ogfs_enable_stats=False, grad_mode=grad_mode)
    K, color, pid, wpix, pw, radii, geom, binning, img, _ = out
    gr = ext.rasterize_gaussians_backward(_t(bg), _t(g["means3D"]), radii, col, _t(g["scales"]),
                                          _t(g["rotations"]), 1.0, torch.Tensor([]), vm, pm,
                                          cam.tanfovx, cam.tanfovy, _t(dL), sh, deg, cp, geom, K,
                                          binning, img, False, grad_mode=grad_mode,
                                          opacities=_t(g["opacities"].reshape(P)))
    return out, gr


def _rel(a, b):
    a = a.double().cpu()
    b = b.double().cpu()
    return float((a - b).norm() / b.norm().clamp_min(1e-30))


@pytest.fixture(params=[0, 1], ids=["layout0", "layout1"])
def bwd_layout(request):
    """Both backward tile layouts (four waves / one wave per tile)."""
    from filtergs_amd import _lib
    before = _lib.get_backward_layout()
    _lib.set_backward_layout(request.param)
    yield request.param
    _lib.set_backward_layout(before)


@pytest.fixture(params=[0, 1], ids=["fwd0", "fwd1"])
def fwd_layout(request):
    """Both forward tile layouts."""
    from filtergs_amd import _lib
    before = _lib.get_forward_layout()
    _lib.set_forward_layout(request.param)
    yield request.param
    _lib.set_forward_layout(before)


@pytest.mark.parametrize("path", ["shs", "colors"])
def test_batch_matches_per_view(path, bwd_layout, fwd_layout):
    g, _ = small_scene(n=20000, seed=171, dist=5.0)
    cams = _cams()
    P = g["means3D"].shape[0]
    bgs = [(0.1, 0.2, 0.3), (0.0, 0.0, 0.0), (1.0, 0.5, 0.25), (0.3, 0.3, 0.3)]
    rng = np.random.default_rng(9)
    dLs = [rng.normal(size=(3, c.height, c.width)).astype(np.float32) for c in cams]
    colors = None
    if path == "colors":
        colors = [rng.uniform(0, 1, size=(P, 3)).astype(np.float32) for _ in cams]
    singles = [_single(g, c, bg, dL, None if colors is None else colors[v])
               for v, (c, bg, dL) in enumerate(zip(cams, bgs, dLs))]

    br = BatchRasterizer(cams, P, bg=bgs)
    if path == "shs":
        br.bind(_t(g["means3D"]), _t(g["opacities"]), _t(g["scales"]), _t(g["rotations"]),
                shs=_t(g["shs"]))
    else:
        br.bind(_t(g["means3D"]), _t(g["opacities"]), _t(g["scales"]), _t(g["rotations"]),
                colors_precomp=_t(np.stack(colors)))
    K = br.forward()
    assert K == sum(int(s[0][0]) for s in singles)
    assert br.K_cap >= K
    for v, (out, _) in enumerate(singles):
        assert torch.equal(br.radii[v], out[5]), v
        assert torch.equal(br.view_color(v), out[1]), v
        assert torch.equal(br.view_pixels(br.point_id, v), out[2]), v
        assert torch.equal(br.view_pixels(br.weight_pixel, v), out[3]), v
        assert torch.equal(br.point_weight[v], out[4]), v
    grads = br.backward([_t(d) for d in dLs])
    for v, (_, gr) in enumerate(singles):
        gm2, gcol = br.view_grads(v)
        assert _rel(gm2, gr[0]) <= 1e-5, v
        assert _rel(gcol, gr[1]) <= 1e-5, v
    summed = {k: sum(s[1][i] for s in singles) for k, i in
              (("opacities", 2), ("means3D", 3), ("cov3D", 4), ("shs", 5), ("scales", 6),
               ("rotations", 7))}
    for k in ("opacities", "means3D", "cov3D", "scales", "rotations") + (("shs",) if path == "shs" else ()):
        got = grads[k].reshape(summed[k].shape)
        assert _rel(got, summed[k]) <= 1e-4, (k, _rel(got, summed[k]))


def test_batch_graph_replay_and_capacity():
    g, _ = small_scene(n=15000, seed=181, dist=5.0)
    cams = _cams()[:3]
    P = g["means3D"].shape[0]
    br = BatchRasterizer(cams, P, bg=(0.2, 0.1, 0.0))
    means = _t(g["means3D"])
    br.bind(means, _t(g["opacities"]), _t(g["scales"]), _t(g["rotations"]), shs=_t(g["shs"]))
    assert br.K_cap == 0
    K = br.forward()          # grows the capacity from zero
    assert br.K_cap >= K > 0
    rng = np.random.default_rng(10)
    dL = _t(rng.normal(size=(3 * int(br.sizes.pixels),)).astype(np.float32))
    eager = br.backward(dL)
    ref_color = br.color.clone()
    ref = {k: (v.clone() if v is not None else None) for k, v in eager.items()}
    graph, out = br.capture(dL)
    for _ in range(3):
        br.color.zero_()
        graph.replay()
        torch.cuda.synchronize()
        assert int(br.counts[1]) == K
        assert torch.equal(br.color, ref_color)
        for k, v in ref.items():
            if v is not None:
                assert _rel(out[k], v) <= 1e-5, k
        # eager work between replays
        torch.cuda.synchronize()
        (means * 1.0).sum().item()
    # moving the Gaussians changes what a replay renders
    means.add_(0.01)
    graph.replay()
    torch.cuda.synchronize()
    assert not torch.equal(br.color, ref_color)


@pytest.mark.parametrize("V", [1, 20, 64])
def test_batch_view_counts(V):
    """One view, more views than a 4096-key sort chunk holds per wave (the
    depth sort reads the preprocess's per-wave segments of 64 * V keys) and
    the most views a batch takes (the backward's per-Gaussian masks of
    visible views are 64 bits): the batch equals the per-view renders, with
    a Gaussian count that leaves the last wave partial."""
    g, _ = small_scene(n=5037, seed=191, dist=5.0)
    base = _cams()
    cams = [base[v % 4] if v < 4 else make_camera(64 + 8 * v, 48 + 4 * v,
                                                    center=(4.0 * np.cos(v), 4.0 * np.sin(v), 0.5 * (v % 3)),
                                                    fovx_deg=50.0 + v)
            for v in range(V)]
    if V > 1:
        # a view that sees nothing: every (view, Gaussian) key of it is culled
        cams[V // 2] = make_camera(80, 60, center=(0.0, -5.0, 0.5), target=(0.0, -10.0, 0.5), fovx_deg=50.0)
    P = g["means3D"].shape[0]
    rng = np.random.default_rng(V)
    dLs = [rng.normal(size=(3, c.height, c.width)).astype(np.float32) for c in cams]
    singles = [_single(g, c, (0.1, 0.2, 0.3), dL) for c, dL in zip(cams, dLs)]
    br = BatchRasterizer(cams, P, bg=(0.1, 0.2, 0.3))
    br.bind(_t(g["means3D"]), _t(g["opacities"]), _t(g["scales"]), _t(g["rotations"]), shs=_t(g["shs"]))
    K = br.forward()
    assert K == sum(int(s[0][0]) for s in singles)
    if V > 1:
        assert int(singles[V // 2][0][0]) == 0 and not br.radii[V // 2].any()
    for v, (out, _) in enumerate(singles):
        assert torch.equal(br.radii[v], out[5]), v
        assert torch.equal(br.view_color(v), out[1]), v
        assert torch.equal(br.view_pixels(br.point_id, v), out[2]), v
    grads = br.backward([_t(d) for d in dLs])
    for k, idx in (("opacities", 2), ("means3D", 3), ("cov3D", 4), ("shs", 5), ("scales", 6),
                   ("rotations", 7)):
        summed = sum(s[1][idx] for s in singles)
        assert _rel(grads[k].reshape(summed.shape), summed) <= 1e-4, k


@pytest.mark.parametrize("deg", [0, 1, 2])
def test_batch_sh_degree_below_storage(deg):
    """Rows of 16 SH coefficients rendered at a lower active degree: the
    batched backward's SH pass reads the first (deg + 1)^2 coefficients and
    gives the rest zero gradient, like the per-view path."""
    g, _ = small_scene(n=9000, seed=201 + deg, dist=5.0)
    cams = _cams()[:3]
    P = g["means3D"].shape[0]
    rng = np.random.default_rng(30 + deg)
    dLs = [rng.normal(size=(3, c.height, c.width)).astype(np.float32) for c in cams]
    singles = [_single(g, c, (0.1, 0.2, 0.3), dL, deg=deg) for c, dL in zip(cams, dLs)]
    br = BatchRasterizer(cams, P, bg=(0.1, 0.2, 0.3), sh_degree=deg)
    br.bind(_t(g["means3D"]), _t(g["opacities"]), _t(g["scales"]), _t(g["rotations"]), shs=_t(g["shs"]))
    K = br.forward()
    assert K == sum(int(s[0][0]) for s in singles)
    for v, (out, _) in enumerate(singles):
        assert torch.equal(br.view_color(v), out[1]), v
    grads = br.backward([_t(d) for d in dLs])
    nact = (deg + 1) ** 2
    assert not grads["shs"].reshape(P, -1, 3)[:, nact:].any()
    for k, idx in (("means3D", 3), ("cov3D", 4), ("shs", 5), ("opacities", 2)):
        summed = sum(s[1][idx] for s in singles)
        assert _rel(grads[k].reshape(summed.shape), summed) <= 1e-4, k


def _rel_np(a, r):
    return float(np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30))


@pytest.mark.parametrize("grad_mode", [0, 1])
@pytest.mark.parametrize("scale_modifier", [0.5, 1.7])
def test_batch_vs_oracle(orc, grad_mode, scale_modifier):
    """The batched kernels against the oracle directly, per view: radii, the
    preprocess outputs, the sorted kv list and the tile ranges bit-exact
    against the FP32 emulation; images against FP32 and FP64; n_contrib on
    the pixels whose colour agrees; per-view dL/dmean2D, dL/dcolors and the
    parameter gradients summed over the views against FP64.  Both gradient
    modes and scale modifiers off 1, with a tenth of the Gaussians at
    opacity in [0.99, 1) so the alpha clamp is met."""
    from tests._scenes import orc_settings
    g, _ = small_scene(n=12000, seed=231, dist=5.0)
    P = g["means3D"].shape[0]
    rng = np.random.default_rng(232)
    hi = rng.choice(P, P // 10, replace=False)
    g["opacities"][hi, 0] = rng.uniform(0.99, 0.9999, hi.size).astype(np.float32)
    cams = _cams()
    bgs = [(0.1, 0.2, 0.3), (0.0, 0.0, 0.0), (1.0, 0.5, 0.25), (0.3, 0.3, 0.3)]
    dLs = [rng.normal(size=(3, c.height, c.width)).astype(np.float32) for c in cams]
    br = BatchRasterizer(cams, P, bg=bgs, grad_mode=grad_mode, scale_modifier=scale_modifier)
    br.bind(_t(g["means3D"]), _t(g["opacities"]), _t(g["scales"]), _t(g["rotations"]), shs=_t(g["shs"]))
    K = br.forward()
    assert 0 < K <= br.K_cap
    grads = br.backward([_t(d) for d in dLs])
    names = ("dL_dopacity", "dL_dmeans3D", "dL_dsh", "dL_dscales", "dL_drot")
    keys = ("opacities", "means3D", "shs", "scales", "rotations")
    sums = {m: {k: 0.0 for k in names} for m in ("f64", "f32")}
    Ksum = 0
    for v, (cam, bg, dL) in enumerate(zip(cams, bgs, dLs)):
        s = orc_settings(orc, cam, sh_degree=3, bg=bg, grad_mode=grad_mode,
                         scale_modifier=scale_modifier)
        o = {}
        for mode, dt in (("f32", np.float32), ("f64", np.float64)):
            d = {k: g[k].astype(dt) for k in g}
            f = orc.render(s, d["means3D"], d["opacities"], d["scales"], d["rotations"], d["shs"],
                           mode=mode, nthreads=8, decision_hash=True)
            b = orc.backward(s, f, dL.astype(dt), d["means3D"], d["opacities"], d["scales"],
                             d["rotations"], d["shs"], mode=mode)
            o[mode] = (f, b)
            for k in names:
                sums[mode][k] = sums[mode][k] + b[k].reshape(P, -1)
        f32, f64 = o["f32"][0], o["f64"][0]
        pre = f32["pre32"]
        # integer stage and its FP32 inputs
        assert np.array_equal(_np(br.radii[v]), pre["radii"]), v
        vis = pre["radii"] > 0
        rec = br.view_records(v)
        for k in ("means2D", "conic_opacity"):
            assert np.array_equal(_np(rec[k])[vis].view(np.uint32), pre[k][vis].view(np.uint32)), (v, k)
        tile, gid, ranges = br.view_sorted(v, K)
        okeys = f32["keys"]
        assert np.array_equal(_np(tile).astype(np.uint64), okeys >> np.uint64(32)), v
        assert np.array_equal(_np(gid).astype(np.uint32), f32["vals"]), v
        assert np.array_equal(_np(ranges).astype(np.uint32), f32["ranges"]), v
        Ksum += len(f32["vals"])
        # image
        img = _np(br.view_color(v))
        d32 = np.abs(img - f32["color"]).max(0)
        agree = d32 <= 2e-5
        assert (~agree).sum() <= 10 + 1e-5 * agree.size and d32.max() <= 0.01, v
        same = agree & (f64["decision_hash"] == f32["decision_hash"])
        assert np.abs(img - f64["color"]).max(0)[same].max() <= 1e-4, v
        assert (~same).sum() <= 1e-3 * same.size, v
        fT, nc = br.view_state(v)
        assert (_np(nc).astype(np.uint32)[agree] != f32["n_contrib"][agree]).sum() <= 2, v
        oT = f32["final_T"]
        assert (np.abs(_np(fT) - oT)[agree] <= 1e-4 * oT[agree] + 1e-7).mean() > 0.999, v
        # per-view screen-space gradients
        b64 = o["f64"][1]
        b32 = o["f32"][1]
        for k, t in zip(("dL_dmean2D", "dL_dcolors"), br.view_grads(v)):
            a = _np(t).astype(np.float64)
            e, e32 = _rel_np(a, b64[k]), _rel_np(b32[k].astype(np.float64), b64[k])
            assert e <= max(1e-4, 3 * e32 + 1e-5), (v, k, e, e32)
    assert K == Ksum
    for k, name in zip(keys, names):
        a = _np(grads[k]).astype(np.float64).reshape(P, -1)
        r = sums["f64"][name]
        r32 = sums["f32"][name].astype(np.float64)
        e, e32 = _rel_np(a, r), _rel_np(r32, r)
        assert e <= max(1e-4, 3 * e32 + 1e-5), (k, e, e32)
        # the batch sums dL/dcov3D over the views before the scale / rotation
        # chain, the oracle after it: equal up to rounding, which the chain's
        # cancellations (sub-pixel Gaussians at scale_modifier 0.5) amplify on
        # a few elements
        dev = np.abs(a - r)
        scale = np.abs(r).max()
        far = dev > 2e-3 * scale + 1e-9
        assert far.sum() <= max(3, 1e-4 * far.size), (k, int(far.sum()))
        assert dev.max() <= 1e-2 * scale + 1e-9, (k, float(dev.max()), float(scale))


def _np(t):
    return t.detach().cpu().numpy()


def test_batch_tile_order_does_not_change_outputs():
    """The tile order of the batched one-wave blend kernels moves work
    between L2s and waves only: the XCD run length (0 = launch order) and the
    persistent launches over the cost-ordered queue (several run lengths, on
    more tiles than the persistent grid holds waves, so waves claim many
    tiles) give the same image, point-id and weight bits, and gradients equal
    up to float summation order.  The outputs are poisoned before each run,
    so a tile the queue skipped or worked twice would show."""
    from filtergs_amd import _lib
    g, _ = small_scene(n=20000, seed=241, dist=5.0)
    base = _cams()
    cams = [make_camera(512 + 32 * v, 288 + 16 * v, center=(4.0 * np.cos(v), 4.0 * np.sin(v), 0.5),
                        fovx_deg=55.0) for v in range(12)] + base
    P = g["means3D"].shape[0]
    br = BatchRasterizer(cams, P, bg=(0.1, 0.2, 0.3))
    assert br.sizes.tiles > 10000
    br.bind(_t(g["means3D"]), _t(g["opacities"]), _t(g["scales"]), _t(g["rotations"]), shs=_t(g["shs"]))
    dL = _t(np.random.default_rng(242).normal(size=(3 * int(br.sizes.pixels),)).astype(np.float32))
    res = []
    _lib.set_forward_layout(1)      # the one-wave forward at any tile count
    _lib.set_backward_layout(1)
    try:
        for chunk, queue, by_cost in ((0, 0, 1), (1, 0, 1), (32, 0, 1), (5, 0, 1), (32, 1, 1), (32, 32, 1),
                                      (0, 7, 1), (0, 4096, 1), (0, 1, 0), (0, 13, 0)):
            _lib.set_xcd_chunk(chunk)
            _lib.set_tile_queue(queue, by_cost)
            br.color.fill_(float("nan"))
            br.point_id.fill_(-7)
            br.weight_pixel.fill_(-1.0)
            br.forward()
            gr = br.backward(dL)
            res.append((br.color.clone(), br.point_id.clone(), br.point_weight.clone(),
                        br.weight_pixel.clone(), {k: v.clone() for k, v in gr.items() if v is not None}))
    finally:
        _lib.set_xcd_chunk(32)
        _lib.set_tile_queue(0)
        _lib.set_forward_layout(2)
    assert not torch.isnan(res[0][0]).any() and int((res[0][1] == -7).sum()) == 0
    for r in res[1:]:
        for a, b in zip(r[:4], res[0][:4]):
            assert torch.equal(a, b)
        for k in r[4]:
            assert _rel(r[4][k], res[0][4][k]) <= 1e-4, k
