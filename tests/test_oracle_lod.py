"""Pins the oracle's LoD cut (FilterGS/cuda/compute_radius_kernel_v2.cu:231-353)
with a recursive definition over parent pointers on random trees, the tree
format invariants of apps/ancestor.py:21-81, and a closed-form radius."""
import math

import numpy as np

from filtergs_amd.camera import make_camera


def random_tree(n_roots=50, depth=5, max_child=4, seed=0, spread=3.0):
    """Random LoD tree: child scale = parent/2, children jittered around the
    parent.  Returns arrays plus the parent pointers."""
    rng = np.random.default_rng(seed)
    means, logs, parent, level = [], [], [], []
    frontier = []
    for _ in range(n_roots):
        means.append(rng.uniform(-spread, spread, 3))
        logs.append(np.full(3, math.log(0.4)) + rng.normal(0, 0.1, 3))
        parent.append(-1)
        level.append(0)
        frontier.append(len(means) - 1)
    for d in range(1, depth):
        nxt = []
        for p in frontier:
            nc = int(rng.integers(0, max_child + 1)) if d > 1 else int(rng.integers(1, max_child + 1))
            for _ in range(nc):
                means.append(means[p] + rng.normal(0, np.exp(logs[p]).mean(), 3))
                logs.append(logs[p] - math.log(2.0) + rng.normal(0, 0.05, 3))
                parent.append(p)
                level.append(d)
                nxt.append(len(means) - 1)
        frontier = nxt
    N = len(means)
    parent = np.array(parent)
    level = np.array(level)
    has_child = np.zeros(N, bool)
    has_child[parent[parent >= 0]] = True
    node_index = np.where(has_child, np.arange(N), -1).astype(np.int32)
    L = int(level.max()) + 1
    anc = np.full((N, L), -1, np.int32)
    for i in range(N):
        a = parent[i]
        while a >= 0:
            anc[i, level[a]] = a
            a = parent[a]
    q = rng.normal(size=(N, 4))
    return dict(means3D=np.array(means, np.float32), log_scales=np.array(logs, np.float32),
                rotations=q.astype(np.float32), node_index=node_index, ancestor_path=anc,
                parent=parent, level=level)


def test_tree_format_invariants():
    t = random_tree(seed=1)
    anc, level, parent = t["ancestor_path"], t["level"], t["parent"]
    for i in range(len(level)):
        row = anc[i]
        assert (row >= 0).sum() == level[i]
        assert (row[:level[i]] >= 0).all() and (row[level[i]:] == -1).all()
        if level[i] > 0:
            assert row[level[i] - 1] == parent[i]


def _camera():
    return make_camera(640, 480, center=(0.0, -9.0, 2.0), fovx_deg=70.0)


def _brute_mask(t, r2d, vis, rmax, thr):
    """Recursive definition: a leaf is dropped when any internal ancestor
    (following parent pointers) is visible with 0 < r <= rmax; an internal
    node is kept when visible with r <= rmax; then the threshold rule."""
    N = len(t["parent"])
    internal = t["node_index"] != -1
    out = np.zeros(N, bool)
    for i in range(N):
        if not vis[i]:
            continue
        if internal[i]:
            keep = r2d[i] <= rmax
        else:
            keep = True
            a = t["parent"][i]
            while a >= 0:
                if internal[a] and vis[a] and 0 < r2d[a] <= rmax:
                    keep = False
                a = t["parent"][a]
        if keep and thr > 0 and r2d[i] > thr:
            keep = False
        out[i] = keep
    return out


def test_lod_cut_matches_recursive_definition(orc):
    t = random_tree(n_roots=120, depth=6, seed=2)
    cam = _camera()
    fx = cam.width / (2 * cam.tanfovx)
    fy = cam.height / (2 * cam.tanfovy)
    for rmax, thr, pad in ((5.0, 700.0, 0.1), (1.5, 30.0, 0.0), (50.0, 0.0, 0.3)):
        mask, r2d, depth = orc.lod_cut(t["means3D"], t["log_scales"], t["rotations"],
                                       cam.full_proj_transform, cam.world_view_transform,
                                       t["node_index"], t["ancestor_path"], fx, fy, cam.tanfovx,
                                       cam.tanfovy, pad, rmax, thr)
        # visibility from an independent NDC computation
        X = np.concatenate([t["means3D"].astype(np.float64), np.ones((len(depth), 1))], 1)
        h = X @ cam.full_proj_transform.astype(np.float64)
        ndc = h[:, :3] / (h[:, 3:] + 1e-7)
        vis = (ndc[:, 2] > 0) & (ndc[:, 2] < 1) & (np.abs(ndc[:, 0]) < 1 + pad) & (np.abs(ndc[:, 1]) < 1 + pad)
        border = (np.abs(np.abs(ndc[:, 0]) - (1 + pad)) < 1e-5) | (np.abs(np.abs(ndc[:, 1]) - (1 + pad)) < 1e-5)
        assert np.array_equal(vis[~border], (r2d > 0)[~border] | (vis & (r2d == 0))[~border])
        np.testing.assert_allclose(depth, ndc[:, 2], atol=1e-5)
        expect = _brute_mask(t, r2d, r2d > 0, rmax, thr)
        assert np.array_equal(mask, expect)
        assert 0 < mask.sum() < len(mask)


def test_lod_radius_closed_form(orc):
    """Isotropic node on the optical axis at distance z: Sigma2D = (f s / z)^2 I,
    floored at 0.3, disc clamp 0.1: r = 3 sqrt(a + sqrt(0.1))."""
    cam = make_camera(400, 400, center=(0.0, -10.0, 0.0), fovx_deg=60.0)
    f = cam.width / (2 * cam.tanfovx)
    zs = np.array([3.0, 5.0, 8.0])
    s = np.array([0.05, 0.01, 0.3])
    means = np.stack([np.zeros(3), zs - 10.0, np.zeros(3)], 1).astype(np.float32)
    mask, r2d, _ = orc.lod_cut(means, np.log(np.repeat(s[:, None], 3, 1)).astype(np.float32),
                               np.tile([1.0, 0, 0, 0], (3, 1)).astype(np.float32),
                               cam.full_proj_transform, cam.world_view_transform,
                               np.array([5, 6, 7], np.int32), np.full((3, 1), -1, np.int32), f, f,
                               cam.tanfovx, cam.tanfovy, 0.1, 1000.0, 0.0)
    a = np.maximum((f * s / zs) ** 2, 0.3)
    np.testing.assert_allclose(r2d, 3 * np.sqrt(a + math.sqrt(0.1)), rtol=1e-5)
    assert mask.all()
