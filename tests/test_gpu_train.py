"""Training-side kernels on the GPU: SparseOptimizer.step against the oracle's
restatement of the reference Adam and against the reference's own torch
operation sequence run on the device; Counter.update_by_output against the
oracle (integer counters exact)."""
import math

import numpy as np
import pytest
import torch
import torch.nn as nn

from filtergs_amd.train import Counter, SparseOptimizer
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu
DEV = "cuda"


class _Model(nn.Module):
    def __init__(self, N, seed):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        self.xyz = nn.Parameter(torch.randn((N, 3), generator=g).to(DEV))
        self.scaling = nn.Parameter(torch.randn((N, 3), generator=g).to(DEV))
        self.colors = nn.Parameter(torch.randn((N, 3), generator=g).to(DEV))
        self.opacity = nn.Parameter(torch.randn((N, 1), generator=g).to(DEV))


def _torch_reference_adam(param, grad, m, v, step, lr, eps=1e-15, b1=0.9, b2=0.999):
    """The reference's _single_tensor_adam operation sequence, on the device."""
    m.mul_(b1).add_(grad, alpha=1 - b1)
    v.mul_(b2).addcmul_(grad, grad.conj(), value=1 - b2)
    bc1 = 1 - b1 ** step
    bc2 = 1 - b2 ** step
    denom = (v.sqrt() / math.sqrt(bc2)).add_(eps)
    param.add_(-(lr / bc1) * (m / denom))
    return param, m, v


@pytest.mark.parametrize("amsgrad", [False, True])
def test_sparse_adam_matches_oracle_and_torch(amsgrad):
    N = 5000
    model = _Model(N, seed=1)
    lr_dict = {"xyz": 1.6e-4, "xyz_final": 1.6e-6, "max_steps": 30000, "scaling": 5e-3,
               "colors": 2.5e-3, "opacity": 5e-2}
    opt = SparseOptimizer(["xyz", "scaling", "colors", "opacity"], lr_dict, model, xyz_scale=1.0,
                          use_amsgrad=amsgrad)
    rng = np.random.default_rng(2)
    state = {k: (np.zeros((N, getattr(model, k).shape[1]), np.float32),) * 2 + (np.zeros((N, getattr(model, k).shape[1]), np.float32),)
             for k in ["xyz", "scaling", "colors", "opacity"]}
    ref = {k: getattr(model, k).detach().cpu().numpy().copy() for k in state}
    tref = {k: (getattr(model, k).detach().clone(), torch.zeros_like(getattr(model, k)),
                torch.zeros_like(getattr(model, k))) for k in state}
    for it in range(1, 4):
        index = torch.as_tensor(rng.choice(N, 1200, replace=False), device=DEV)
        flag = torch.as_tensor(rng.random(1200) < 0.8, device=DEV)
        params = {k: nn.Parameter(getattr(model, k).detach()[index].clone()) for k in state}
        grads = {k: torch.as_tensor(rng.normal(size=tuple(params[k].shape)).astype(np.float32), device=DEV)
                 for k in state}
        for k in state:
            params[k].grad = grads[k]
        opt.step(model, index, params, flag)
        step = float(it)
        for k in state:
            lr = opt.xyz_scheduler_args(step) if k == "xyz" else (
                opt.scaling_scheduler_args(step) if k == "scaling" else lr_dict[k])
            p_sel = params[k].detach().cpu().numpy()
            m0, v0, vm0 = state[k]
            ref[k], m1, v1, vm1 = orc.sparse_adam_step(ref[k], p_sel, grads[k].cpu().numpy(), m0, v0,
                                                        index.cpu().numpy(), flag.cpu().numpy(), lr, step,
                                                        max_exp_avg_sq=vm0 if amsgrad else None)
            state[k] = (m1, v1, vm1 if amsgrad else vm0)
            np.testing.assert_allclose(getattr(model, k).detach().cpu().numpy(), ref[k], rtol=2e-6, atol=1e-7,
                                       err_msg=k)
            np.testing.assert_allclose(opt.exp_avg[k].cpu().numpy(), m1, rtol=2e-6, atol=1e-9)
            np.testing.assert_allclose(opt.exp_avg_sq[k].cpu().numpy(), v1, rtol=2e-6, atol=1e-12)
            if not amsgrad:
                # the reference's torch sequence on the device, selected rows only
                tp, tm, tv = tref[k]
                sel = index[flag]
                p_new, m_new, v_new = _torch_reference_adam(params[k].detach()[flag].clone(), grads[k][flag],
                                                            tm[sel].clone(), tv[sel].clone(), step, lr)
                tp[sel], tm[sel], tv[sel] = p_new, m_new, v_new
                torch.testing.assert_close(getattr(model, k).detach()[sel], tp[sel], rtol=2e-6, atol=1e-7)


def test_sparse_adam_bad_index():
    model = _Model(10, seed=3)
    opt = SparseOptimizer(["colors"], {"colors": 1e-3}, model)
    index = torch.tensor([1, 2, 99], device=DEV)
    params = {"colors": nn.Parameter(torch.zeros((3, 3), device=DEV))}
    params["colors"].grad = torch.ones((3, 3), device=DEV)
    with pytest.raises(IndexError):
        opt.step(model, index, params, torch.ones(3, dtype=torch.bool, device=DEV))


def test_counter_update_matches_oracle():
    N = 20000
    ctr = Counter(N).to(DEV)
    rng = np.random.default_rng(5)
    ctr.weights_max.copy_(torch.as_tensor(rng.random(N).astype(np.float32)))
    ctr.radii_max.copy_(torch.as_tensor(rng.integers(0, 20, N).astype(np.int16)))
    ctr.radii_max_max.copy_(torch.as_tensor(rng.integers(0, 50, N).astype(np.int32)))
    ref = {k: getattr(ctr, k).cpu().numpy().copy() for k in
           ("create_steps", "visible_count", "weights_max", "weights_sum", "radii_max", "radii_max_max",
            "area_sum", "grad_sum")}
    outputs = {"render": [], "visibility_flag": [], "radii": [], "point_id": [], "point_count": [],
               "point_weight": [], "viewspace_points": []}
    for _ in range(3):
        n = 6000
        vis = rng.choice(N, n, replace=False)
        leaf, node = vis[:5000], vis[5000:]
        radii = np.where(rng.random(n) < 0.7, rng.integers(1, 40, n), 0).astype(np.int32)
        w = rng.random(n).astype(np.float32)
        pid = np.sort(rng.choice(n, 900, replace=False))
        pcount = rng.integers(1, 300, 900)
        grad = rng.normal(size=(n, 3)).astype(np.float32)
        vp = torch.zeros((n, 3), device=DEV, requires_grad=True)
        vp.grad = torch.as_tensor(grad, device=DEV)
        outputs["render"].append(torch.zeros((3, 4, 4), device=DEV))
        outputs["visibility_flag"].append({"index": torch.as_tensor(leaf, device=DEV),
                                           "index_node": torch.as_tensor(node, device=DEV)})
        outputs["radii"].append(torch.as_tensor(radii, device=DEV))
        outputs["point_id"].append(torch.as_tensor(pid, device=DEV))
        outputs["point_count"].append(torch.as_tensor(pcount, device=DEV))
        outputs["point_weight"].append(torch.as_tensor(w, device=DEV))
        outputs["viewspace_points"].append(vp)
        ref = orc.counter_update(ref, vis, radii, w, pid, pcount, grad)
    ctr.update_by_output(outputs)
    for k, v in ref.items():
        got = getattr(ctr, k).cpu().numpy()
        if v.dtype.kind in "iu":
            assert np.array_equal(got, v), k
        else:
            np.testing.assert_allclose(got, v, rtol=1e-6, atol=1e-6, err_msg=k)
    assert "flag_vis" in outputs["visibility_flag"][0]


def test_render_to_check_equals_rasterizer_point_weight():
    from filtergs_amd.camera import make_camera
    from filtergs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer
    from filtergs_amd.train import render_to_check
    from tests._scenes import small_scene
    g, cam = small_scene(n=4000, W=200, H=150, seed=12)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32, device=DEV)
    rs = GaussianRasterizationSettings(
        image_height=cam.height, image_width=cam.width, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy,
        bg=t([0.0, 0.0, 0.0]), scale_modifier=1.0, viewmatrix=t(cam.world_view_transform),
        projmatrix=t(cam.full_proj_transform), sh_degree=0, campos=t(cam.center), prefiltered=False,
        debug=False)
    r = GaussianRasterizer(rs)
    xyz, sc, rot, op = t(g["means3D"]), t(g["scales"]), t(g["rotations"]), t(g["opacities"].reshape(-1, 1))
    pw = render_to_check(r, xyz, sc, rot, op)
    ref = r(means3D=xyz, means2D=torch.zeros_like(xyz), shs=None, colors_precomp=torch.ones_like(xyz),
            opacities=op, scales=sc, rotations=rot, cov3D_precomp=None)[4]
    assert torch.equal(pw, ref) and float(pw.max()) > 0


def _act_data(n, deg, seed):
    rng = np.random.default_rng(seed)
    K = max(3, (deg + 1) ** 2 - 1) + 1
    d = dict(xyz=rng.normal(0, 2, (n, 3)), scaling=rng.normal(-1, 0.5, (n, 3)), rotation=rng.normal(0, 1, (n, 4)),
             opacity=rng.normal(0, 2, (n, 1)), colors=rng.normal(0, 0.5, (n, 3)),
             shs=rng.normal(0, 0.1, (n, K, 3)))
    d["rotation"][:3] = 0.0
    return {k: v.astype(np.float32) for k, v in d.items()}


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
def test_activate_rows_forward_backward_vs_oracle(deg):
    """The fused activation of LoG's training gather (activation.py:27-43):
    outputs against the oracle forward, raw-row gradients against the oracle
    backward (float64), with zero quaternions among the rows."""
    from filtergs_amd.train import activate_root_return
    n = 20000
    d = _act_data(n, deg, 40 + deg)
    campos = np.array([0.5, -8.0, 1.5], np.float32)
    leaves = {k: nn.Parameter(torch.as_tensor(v, device=DEV)) for k, v in d.items()}
    camera = {"camera_center": torch.as_tensor(campos, device=DEV)}
    out = activate_root_return(leaves, camera, deg)
    assert out["xyz"] is leaves["xyz"]
    rng = np.random.default_rng(7)
    g = {k: rng.normal(0, 1, tuple(out[k].shape)).astype(np.float32)
         for k in ("scaling", "rotation", "opacity", "colors")}
    torch.autograd.backward([out[k] for k in g], [torch.as_tensor(v, device=DEV) for v in g.values()])
    K = d["shs"].shape[1]
    ref = orc.activate_gather(np.arange(n), d["xyz"], d["scaling"], d["rotation"], d["opacity"].reshape(-1),
                              d["colors"], d["shs"] if deg else None, deg, campos)
    for k in ("scaling", "rotation", "colors"):
        np.testing.assert_allclose(out[k].detach().cpu().numpy(), ref[k], rtol=2e-5, atol=2e-6)
    np.testing.assert_allclose(out["opacity"].detach().cpu().numpy().reshape(-1), ref["opacity"],
                               rtol=2e-5, atol=2e-6)
    bw = orc.activate_backward(d["xyz"], d["scaling"], d["rotation"], d["opacity"], K, deg, campos,
                               g["scaling"], g["rotation"], g["opacity"], g["colors"])
    pairs = (("scaling", "log_scaling"), ("rotation", "rotation"), ("opacity", "opacity_logit"),
             ("colors", "features_dc"))
    for leaf, key in pairs:
        got = leaves[leaf].grad.detach().cpu().numpy().reshape(bw[key].shape)
        want = bw[key]
        # rows 0-2 hold zero quaternions (rotation gradient g / 1e-12): own scale
        for rows in (slice(0, 3), slice(3, None)):
            scale = np.abs(want[rows]).max()
            err = np.abs(got[rows] - want[rows]).max()
            assert err <= 2e-5 * scale, (leaf, rows, err, scale)
    if deg == 0:
        assert leaves["shs"].grad is None
    else:
        got = leaves["shs"].grad.detach().cpu().numpy()
        assert np.abs(got - bw["features_rest"]).max() <= 2e-5 * np.abs(bw["features_rest"]).max()
        assert not got[:, (deg + 1) ** 2 - 1:].any()
    assert leaves["xyz"].grad is None


def _torch_log_activation(ret, campos, deg):
    """The reference's activation sequence in torch (activation.py:27-43,
    sh_utils.py:31-72), for the drop-in comparison."""
    C0, C1 = 0.28209479177387814, 0.4886025119029199
    C2 = [1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
          0.5462742152960396]
    colors = ret["colors"] * C0 + 0.5
    if deg > 0:
        d = ret["xyz"].detach() - campos[None]
        d = d / torch.norm(d, dim=-1, keepdim=True)
        x, y, z = d[..., 0:1], d[..., 1:2], d[..., 2:3]
        sh = ret["shs"]
        r = -C1 * y * sh[..., 0, :] + C1 * z * sh[..., 1, :] - C1 * x * sh[..., 2, :]
        if deg > 1:
            xx, yy, zz, xy, yz, xz = x * x, y * y, z * z, x * y, y * z, x * z
            r = (r + C2[0] * xy * sh[..., 3, :] + C2[1] * yz * sh[..., 4, :] +
                 C2[2] * (2.0 * zz - xx - yy) * sh[..., 5, :] + C2[3] * xz * sh[..., 6, :] +
                 C2[4] * (xx - yy) * sh[..., 7, :])
        colors = colors + r
    return {"xyz": ret["xyz"], "scaling": torch.exp(ret["scaling"]), "opacity": torch.sigmoid(ret["opacity"]),
            "rotation": torch.nn.functional.normalize(ret["rotation"]), "colors": colors}


@pytest.mark.parametrize("fix_parent", [True, False])
def test_log_get_all_drop_in(fix_parent):
    """get_all(log_model, camera) against LoG.get_all's sequence
    (level_of_gaussian.py:260-294) in torch: same outputs, same gradients on
    the gathered leaves in visibility_flag['params']."""
    from types import SimpleNamespace
    from filtergs_amd.train import get_all
    N, deg = 50000, 2
    d = _act_data(N, deg, 3)
    bufs = {k: torch.as_tensor(v, device=DEV) for k, v in d.items()}
    gen = torch.Generator(device="cpu").manual_seed(5)
    perm = torch.randperm(N, generator=gen)
    index, index_node = perm[:20000].to(DEV), perm[20000:23000].to(DEV)
    campos = torch.tensor([0.5, -8.0, 1.5], device=DEV)
    camera = {"camera_center": campos}

    def model():
        g = SimpleNamespace(items=lambda: iter(bufs.items()), active_sh_degree=deg,
                            visibility_flag={"index": index, "index_node": index_node})
        return SimpleNamespace(gaussian=g, fix_parent=fix_parent, training=True)

    ours_m, ref_m = model(), model()
    ours = get_all(ours_m, camera)
    # the reference's gather, then its activation chain
    vf = ref_m.gaussian.visibility_flag
    if fix_parent:
        leaves = {k: nn.Parameter(v[index]) for k, v in bufs.items()}
        ret = {k: torch.cat([leaves[k], bufs[k][index_node]]) for k in bufs}
    else:
        leaves = {k: nn.Parameter(v[torch.cat([index, index_node])]) for k, v in bufs.items()}
        ret = leaves
    vf["params"] = leaves
    ref = _torch_log_activation(ret, campos, deg)
    gen = torch.Generator(device=DEV).manual_seed(9)
    loss_w = {k: torch.randn(ref[k].shape, device=DEV, generator=gen) for k in ("scaling", "rotation", "opacity", "colors")}
    for out in (ours, ref):
        sum((out[k] * loss_w[k]).sum() for k in loss_w).backward()
    for k in ("xyz", "scaling", "rotation", "opacity", "colors"):
        torch.testing.assert_close(ours[k], ref[k], rtol=1e-5, atol=1e-6)
    ours_leaves = ours_m.gaussian.visibility_flag["params"]
    for k in ("scaling", "rotation", "opacity", "colors", "shs"):
        torch.testing.assert_close(ours_leaves[k].grad, leaves[k].grad, rtol=1e-4, atol=1e-5)
    assert ours_leaves["xyz"].grad is None and leaves["xyz"].grad is None
