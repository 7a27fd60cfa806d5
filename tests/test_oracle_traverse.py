"""The oracle's TensorTree.traverse (LoG/model/tensor_tree.py:133-188) and
LoG radius (LoG/model/level_of_gaussian.py:66-88), pinned by a different
algorithm (depth-first walk, then a sort by (level, root position, slot
path)), closed forms and special cases."""
import math

import numpy as np
import pytest

from filtergs_amd.camera import make_camera
from oracle import pyoracle as orc
from tests._trees import chain_tree, log_tree


def _cam(W=320, H=240):
    return make_camera(W, H, center=(0.0, 0.0, 0.0), target=(0.0, 10.0, 0.0), fovx_deg=60.0,
                       znear=0.05, zfar=100.0)


def _radius_fn(t, cam):
    fx = cam.width / (2 * cam.tanfovx)
    fy = cam.height / (2 * cam.tanfovy)
    return lambda ids: orc.log_radius(ids, t["xyz"], t["scaling"], t["rotation"],
                                      cam.full_proj_transform, cam.world_view_transform, fx, fy,
                                      cam.tanfovx, cam.tanfovy)


def dfs_cut(node_index, tree, roots, radius, min_px, max_depth):
    """Depth-first restatement: every node reached is decided on its own
    radius; a kept node is listed with its level and slot path, a node that
    would go on past the level limit is listed as a limit node.  The
    breadth-first order is the sort of the kept entries by (level, path),
    then the limit nodes by path."""
    limit = min(30, max_depth)
    kept, at_limit = [], []

    def walk(i, level, path):
        small = radius(np.array([i]))[0] < min_px
        if small or node_index[i] == -1:
            kept.append((level, path, i))
        elif level + 1 > limit:
            at_limit.append((path, i))
        else:
            for slot, c in enumerate(tree[node_index[i]]):
                if c != -1:
                    walk(int(c), level + 1, path + (slot,))

    for pos, r in enumerate(roots):
        walk(int(r), 0, (pos,))
    kept.sort(key=lambda e: (e[0], e[1]))
    at_limit.sort(key=lambda e: e[0])
    return np.array([e[2] for e in kept] + [e[1] for e in at_limit], np.int64)


@pytest.mark.parametrize("seed,C,min_px,max_depth", [
    (0, 2, 3.0, 1000), (1, 4, 3.0, 1000), (2, 4, 8.0, 2), (3, 8, 1.5, 1000), (4, 4, 3.0, 0),
    (5, 3, 0.0, 1000), (6, 4, 1e9, 1000)])
def test_traverse_matches_depth_first_restatement(seed, C, min_px, max_depth):
    t = log_tree(n_roots=40, max_child=C, rounds=5, seed=seed)
    cam = _cam()
    rf = _radius_fn(t, cam)
    rng = np.random.default_rng(seed)
    roots = rng.permutation(np.nonzero(t["index_parent"] == -1)[0])[:30]
    got = orc.tree_traverse(t["node_index"], t["tree"], roots, rf, min_px, max_depth)
    want = dfs_cut(t["node_index"], t["tree"], roots, rf, min_px, max_depth)
    assert np.array_equal(got, want)
    assert got.dtype == np.int64
    if min_px >= 1e9 or max_depth == 0:
        # nothing below the roots is visited: the roots, kept ones first
        assert sorted(got.tolist()) == sorted(roots.tolist())
    if min_px == 0.0:
        # nothing is small: every reachable leaf (the tree is shallower than 30)
        assert (t["node_index"][got] == -1).all()


def test_traverse_level_limit_on_a_chain():
    t = chain_tree(36)
    rf = _radius_fn(t, _cam())
    assert (rf(np.arange(36)) > 3.0).all()
    # nothing is small: the walk stops after level 30 and returns the node there
    assert orc.tree_traverse(t["node_index"], t["tree"], [0], rf, 3.0, 1000).tolist() == [30]
    assert orc.tree_traverse(t["node_index"], t["tree"], [0], rf, 3.0, 7).tolist() == [7]
    assert orc.tree_traverse(t["node_index"], t["tree"], [0], rf, 3.0, 0).tolist() == [0]
    assert orc.tree_traverse(t["node_index"], t["tree"], [0], rf, 1e9, 1000).tolist() == [0]
    t34 = chain_tree(30)
    rf = _radius_fn(t34, _cam())
    # depth 29 < 30: the leaf at the end
    assert orc.tree_traverse(t34["node_index"], t34["tree"], [0], rf, 3.0, 1000).tolist() == [29]


def test_traverse_empty_roots():
    t = log_tree(n_roots=8, seed=9)
    rf = _radius_fn(t, _cam())
    assert orc.tree_traverse(t["node_index"], t["tree"], np.zeros(0, np.int64), rf).size == 0


def test_log_radius_closed_form():
    """An isotropic Gaussian of scale s on the optical axis at depth d: the 2D
    covariance is (f s / d)^2 I, its eigenvalue gap is lifted to sqrt(0.1), so
    r = 3 sqrt((f s / d)^2 + sqrt(0.1)); the rotation's length and sign do
    not matter (normalised), and a zero quaternion stays zero through
    F.normalize, whose rotation matrix 1 - 2(...) is the identity: the same
    radius again.  An isotropic Gaussian whose 2D variance is below 0.3 px^2
    takes the floor: r = 3 sqrt(0.3 + sqrt(0.1))."""
    cam = _cam()
    fx = cam.width / (2 * cam.tanfovx)
    d, s = 10.0, 0.05
    xyz = np.array([[0.0, d, 0.0]] * 3, np.float32)
    scaling = np.full((3, 3), math.log(s), np.float32)
    rot = np.array([[2.0, 0, 0, 0], [-0.5, 0.0, 0.0, 0.0], [0, 0, 0, 0]], np.float32)
    r = orc.log_radius([0, 1, 2], xyz, scaling, rot, cam.full_proj_transform,
                       cam.world_view_transform, fx, fx, cam.tanfovx, cam.tanfovy)
    lam = (fx * s / d) ** 2
    want = 3.0 * math.sqrt(max(lam, 0.3) + math.sqrt(0.1))
    assert r[0] == pytest.approx(want, rel=2e-6)
    assert r[1] == r[0]
    assert r[2] == r[0]
    tiny = np.full((1, 3), math.log(0.001), np.float32)
    r0 = orc.log_radius([0], xyz[:1], tiny, rot[:1], cam.full_proj_transform,
                        cam.world_view_transform, fx, fx, cam.tanfovx, cam.tanfovy)
    assert (fx * 0.001 / d) ** 2 < 0.3
    assert r0[0] == pytest.approx(3.0 * math.sqrt(0.3 + math.sqrt(0.1)), rel=2e-6)
    # the same rows through radius_v2 with the activation done in double
    t = log_tree(n_roots=30, rounds=3, seed=4)
    ids = np.arange(t["xyz"].shape[0])
    q = t["rotation"].astype(np.float64)
    qn = q / np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
    r_v2 = orc.radius_v2(t["xyz"], np.exp(t["scaling"].astype(np.float64)).astype(np.float32),
                         qn.astype(np.float32), cam.full_proj_transform, cam.world_view_transform,
                         fx, fx, cam.tanfovx, cam.tanfovy)
    r_log = orc.log_radius(ids, t["xyz"], t["scaling"], t["rotation"], cam.full_proj_transform,
                           cam.world_view_transform, fx, fx, cam.tanfovx, cam.tanfovy)
    np.testing.assert_allclose(r_log, r_v2, rtol=1e-5, atol=1e-6)
    assert (r_log > 0).mean() > 0.5
