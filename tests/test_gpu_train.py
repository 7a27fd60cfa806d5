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
