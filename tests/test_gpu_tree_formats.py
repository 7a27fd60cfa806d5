"""Tree formats on the GPU: the device ancestor-row rebuild against the
oracle's restatement of apps/ancestor.py (bit-exact, with orphans), the LoD
cut over parent links against the cut over ancestor rows and the oracle
(bit-exact masks), and a LoG checkpoint file through the loader into a
FilterGS render."""
import numpy as np
import pytest
import torch

from filtergs_amd import lod
from filtergs_amd.camera import make_camera
from oracle import pyoracle as orc
from tests.test_oracle_lod import random_tree
from tests.test_oracle_tree_formats import random_forest

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _t(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


@pytest.mark.parametrize("on_gpu", [False, True])
def test_rebuild_ancestor_matches_oracle(on_gpu):
    parent, depth = random_forest(11, 20000, seed=9)
    parent = parent.copy()
    rng = np.random.default_rng(2)
    orphans = rng.choice(np.nonzero(depth >= 1)[0], 40, replace=False)
    parent[orphans] = -1
    L = int(depth.max()) + 3
    expect = orc.rebuild_ancestor_path(parent, depth, L)
    dev = DEV if on_gpu else "cpu"
    sd = {"tree.index_parent": torch.as_tensor(parent, dtype=torch.int32, device=dev),
          "tree.depth": torch.as_tensor(depth, dtype=torch.int8, device=dev),
          "tree.ancestor_path": torch.zeros((len(depth), L), dtype=torch.int32, device=dev)}
    rows, max_depth = lod.rebuild_ancestor_path(sd)
    assert rows.device.type == dev and sd["tree.ancestor_path"] is rows
    assert max_depth == int(depth.max())
    assert np.array_equal(rows.cpu().numpy(), expect)
    # no ancestor buffer: width depth.max() + 1
    sd2 = {"tree.index_parent": sd["tree.index_parent"], "tree.depth": sd["tree.depth"]}
    rows2, _ = lod.rebuild_ancestor_path(sd2)
    assert rows2.shape == (len(depth), int(depth.max()) + 1)
    assert np.array_equal(rows2.cpu().numpy(), expect[:, :int(depth.max()) + 1])
    # too narrow
    sd3 = {"tree.index_parent": sd["tree.index_parent"], "tree.depth": sd["tree.depth"],
           "tree.ancestor_path": torch.zeros((len(depth), 2), dtype=torch.int32, device=dev)}
    with pytest.raises(IndexError):
        lod.rebuild_ancestor_path(sd3)


def test_parent_link_cut_equals_row_cut():
    tree = random_tree(n_roots=300, depth=7, seed=13)
    cam = make_camera(800, 600, center=(0.0, -10.0, 2.5), fovx_deg=70.0)
    fx, fy = cam.width / (2 * cam.tanfovx), cam.height / (2 * cam.tanfovy)
    args = (_t(tree["means3D"]), _t(tree["log_scales"]), _t(tree["rotations"]),
            _t(cam.full_proj_transform), _t(cam.world_view_transform))
    ni = _t(tree["node_index"], torch.int32)
    for rmax, thr in ((5.0, 700.0), (1.5, 40.0), (12.0, 0.0)):
        m1, r1, d1 = lod.compute_visibility_radius(*args, ni, _t(tree["ancestor_path"], torch.int32), fx, fy,
                                                   cam.tanfovx, cam.tanfovy, rmax, padding=0.1,
                                                   radius_filter_threshold=thr)
        m2, r2, d2 = lod.compute_visibility_radius_parent(*args, ni, _t(tree["parent"], torch.int32), fx, fy,
                                                          cam.tanfovx, cam.tanfovy, rmax, padding=0.1,
                                                          radius_filter_threshold=thr)
        om, _, _ = orc.lod_cut(tree["means3D"], tree["log_scales"], tree["rotations"],
                               cam.full_proj_transform, cam.world_view_transform, tree["node_index"],
                               tree["ancestor_path"], fx, fy, cam.tanfovx, cam.tanfovy, 0.1, rmax, thr)
        assert np.array_equal(m2.cpu().numpy(), om)
        assert torch.equal(m1, m2) and torch.equal(r1, r2) and torch.equal(d1, d2)
        assert 0 < int(m2.sum()) < len(om)


def test_checkpoint_to_filtergs_render(tmp_path):
    from filtergs_amd import checkpoint
    from filtergs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer
    tree = random_tree(n_roots=200, depth=6, seed=21)
    N = len(tree["level"])
    rng = np.random.default_rng(4)
    sd = {
        "gaussian.xyz": torch.as_tensor(tree["means3D"]),
        "gaussian.scaling": torch.as_tensor(tree["log_scales"]),
        "gaussian.rotation": torch.as_tensor(tree["rotations"]),
        "gaussian.opacity": torch.as_tensor(rng.normal(1.0, 0.5, (N, 1)).astype(np.float32)),
        "gaussian.colors": torch.as_tensor(rng.normal(0, 0.5, (N, 3)).astype(np.float32)),
        "gaussian.shs": torch.as_tensor(rng.normal(0, 0.1, (N, 3, 3)).astype(np.float32)),
        "tree.node_index": torch.as_tensor(tree["node_index"]),
        "tree.index_parent": torch.as_tensor(tree["parent"].astype(np.int32)),
        "tree.depth": torch.as_tensor(tree["level"].astype(np.int8)),
    }
    path = tmp_path / "log_model.pth"
    torch.save({"state_dict": sd, "global_iterations": 7}, str(path))
    loaded = checkpoint.load_statedict(str(path))
    tt = checkpoint.tree_tensors(loaded)            # no ancestor buffer: rebuilt
    assert np.array_equal(tt["ancestor_path"].cpu().numpy(),
                          orc.rebuild_ancestor_path(tree["parent"], tree["level"], int(tree["level"].max()) + 1))
    assert np.array_equal(tt["ancestor_path"].cpu().numpy(), tree["ancestor_path"])
    fg = checkpoint.filtergs_from_statedict(loaded)
    assert fg.max_sh_degree == 1 and fg.active_sh_degree == 1
    cam = make_camera(320, 240, center=(0.0, -9.0, 2.0), fovx_deg=70.0)
    rs = GaussianRasterizationSettings(
        image_height=cam.height, image_width=cam.width, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy,
        bg=_t([0.0, 0.0, 0.0]), scale_modifier=1.0, viewmatrix=_t(cam.world_view_transform),
        projmatrix=_t(cam.full_proj_transform), sh_degree=0, campos=_t(cam.center), prefiltered=False,
        debug=False)
    r = GaussianRasterizer(rs)
    camera = {"camera_center": _t(cam.center), "full_proj_transform": _t(cam.full_proj_transform)}
    n = fg.prepare(r, camera, tree_node_index_gpu=tt["node_index"], ancestor_path_gpu=tt["ancestor_path"],
                   radius_max=5.0, padding=0.1, radius_filter_threshold=700.0)
    assert n > 0
    img = fg.render(camera, r)["render"]
    assert img.shape == (3, cam.height, cam.width) and torch.isfinite(img).all() and float(img.abs().sum()) > 0
