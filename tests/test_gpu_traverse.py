"""TensorTree.traverse on the GPU (LoG/model/tensor_tree.py:133-188) against
the oracle's level loop: the id list must be equal, element for element."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from filtergs_amd._lib import FGSError
from filtergs_amd.rasterizer import GaussianRasterizationSettings
from filtergs_amd.tensor_tree import DeviceTraverse, traverse, tree_traverse
from oracle import pyoracle as orc
from tests._trees import chain_tree, log_tree
from tests.test_oracle_traverse import _cam, _radius_fn

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _t(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), device=DEV, dtype=dtype)


def _gpu(t, cam, roots, min_px=3.0, max_depth=1000, hint=None):
    fx = cam.width / (2 * cam.tanfovx)
    fy = cam.height / (2 * cam.tanfovy)
    return tree_traverse(_t(t["xyz"]), _t(t["scaling"]), _t(t["rotation"]), _t(t["node_index"]),
                         _t(t["tree"]), _t(roots, torch.int64), _t(cam.world_view_transform),
                         _t(cam.full_proj_transform), fx, fy, cam.tanfovx, cam.tanfovy, min_px,
                         max_depth, level_hint=hint)


@pytest.mark.parametrize("seed,C,min_px,max_depth", [
    (0, 2, 3.0, 1000), (1, 4, 3.0, 1000), (2, 4, 8.0, 2), (3, 8, 1.5, 1000), (4, 4, 3.0, 0),
    (5, 3, 0.0, 1000), (6, 4, 1e9, 1000), (7, 4, 3.0, 3)])
def test_traverse_matches_oracle(seed, C, min_px, max_depth):
    print(f"traverse seed={seed} C={C} min_px={min_px} max_depth={max_depth}", flush=True)
    t = log_tree(n_roots=400, max_child=C, rounds=6, seed=seed)
    cam = _cam()
    rng = np.random.default_rng(seed)
    roots = rng.permutation(np.nonzero(t["index_parent"] == -1)[0])[:300]
    want = orc.tree_traverse(t["node_index"], t["tree"], roots, _radius_fn(t, cam), min_px, max_depth)
    for hint in (None, int(t["depth"].max()), 40):
        got = _gpu(t, cam, roots, min_px, max_depth, hint)
        assert got.dtype == torch.int64
        assert np.array_equal(got.cpu().numpy(), want), (hint, got.numel(), want.size)


def test_traverse_level_limit_and_empty():
    cam = _cam()
    t = chain_tree(36)
    for md, want in ((1000, [30]), (7, [7]), (0, [0]), (-5, [0])):
        assert _gpu(t, cam, [0], 3.0, md).tolist() == want
        assert _gpu(t, cam, [0], 3.0, md, hint=35).tolist() == want
    assert _gpu(t, cam, np.zeros(0, np.int64)).numel() == 0
    # a hint below the depth is refused, not answered short
    with pytest.raises(FGSError):
        _gpu(t, cam, [0], 3.0, 1000, hint=10)


def test_traverse_bad_ids():
    cam = _cam()
    t = log_tree(n_roots=50, rounds=3, seed=11)
    N = t["xyz"].shape[0]
    with pytest.raises(FGSError):
        _gpu(t, cam, [0, N])
    bad = dict(t)
    bad["tree"] = t["tree"].copy()
    bad["tree"][0, 0] = N + 5
    bad["node_index"] = t["node_index"].copy()
    roots = np.nonzero(t["index_parent"] == -1)[0]
    with pytest.raises(FGSError):
        _gpu(bad, cam, roots, 0.0)
    # the library still works after the refusals
    want = orc.tree_traverse(t["node_index"], t["tree"], roots, _radius_fn(t, cam))
    assert np.array_equal(_gpu(t, cam, roots).cpu().numpy(), want)


def test_traverse_drop_in_signature():
    """tree.traverse(model, root_index, camera, max_depth) with LoG-shaped
    objects: buffers of a TensorTree, a Gaussian with raw parameters and an
    exp activation, a rasterizer carrying raster_settings."""
    cam = _cam(480, 360)
    t = log_tree(n_roots=200, max_child=4, rounds=6, seed=21)

    class Tree(DeviceTraverse):
        pass

    tree = Tree()
    tree.node_index = _t(t["node_index"])
    tree.tree = _t(t["tree"])
    tree.depth = _t(t["depth"].astype(np.int8))
    tree.min_resolution_pixel = 3
    model = SimpleNamespace(xyz=_t(t["xyz"]), scaling=_t(t["scaling"]), rotation=_t(t["rotation"]),
                            activation=SimpleNamespace(scaling_activation=torch.exp))
    rs = GaussianRasterizationSettings(
        image_height=cam.height, image_width=cam.width, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy,
        bg=_t(np.zeros(3, np.float32)), scale_modifier=1.0, viewmatrix=_t(cam.world_view_transform),
        projmatrix=_t(cam.full_proj_transform), sh_degree=0, campos=_t(cam.center),
        prefiltered=False, debug=False)
    rast = SimpleNamespace(raster_settings=rs)
    roots = np.nonzero(t["index_parent"] == -1)[0]
    for md in (1000, 2):
        got = tree.traverse(model, _t(roots, torch.int64), rast, max_depth=md)
        want = orc.tree_traverse(t["node_index"], t["tree"], roots, _radius_fn(t, cam), 3.0, md)
        assert np.array_equal(got.cpu().numpy(), want)
        assert np.array_equal(traverse(tree, model, _t(roots, torch.int64), rast, md).cpu().numpy(), want)
    # the cached depth bound follows a change of the depth buffer
    tree.depth = _t(np.zeros_like(t["depth"]).astype(np.int8))
    with pytest.raises(FGSError):
        tree.traverse(model, _t(roots, torch.int64), rast)
