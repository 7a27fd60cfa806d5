"""LoG-style Gaussian trees for the traverse tests: the TensorTree layout
(LoG/model/tensor_tree.py:4-26) grown by splits as TensorTree.split does
(:67-97: a split node gets the next row of `tree`, its children the next ids
in slot order), with some children removed afterwards (slot set to -1, a node
left without children becomes a leaf again, :99-127)."""
import math

import numpy as np


def log_tree(n_roots=64, max_child=4, rounds=5, split_frac=0.5, remove_frac=0.1, seed=0,
             spread=4.0, root_scale=0.6, dist=10.0):
    rng = np.random.default_rng(seed)
    xyz = [np.stack([rng.uniform(-spread, spread, n_roots), rng.uniform(dist - spread, dist + spread, n_roots),
                     rng.uniform(-spread, spread, n_roots)], 1)]
    logs = [np.log(root_scale) + rng.normal(0, 0.3, (n_roots, 3))]
    node_index = np.full(n_roots, -1, np.int64)
    depth = np.zeros(n_roots, np.int64)
    parent = np.full(n_roots, -1, np.int64)
    rows = []
    N = n_roots
    for _ in range(rounds):
        leaves = np.nonzero(node_index == -1)[0]
        pick = leaves[rng.random(leaves.size) < split_frac]
        pick = pick[depth[pick] < 30]
        if pick.size == 0:
            break
        cur_xyz = np.concatenate(xyz)
        cur_log = np.concatenate(logs)
        k = pick.size
        node_index[pick] = len(rows) + np.arange(k)
        ids = N + np.arange(k * max_child).reshape(k, max_child)
        rows.extend(list(ids))
        par = np.repeat(pick, max_child)
        sc = np.exp(cur_log[par]).mean(1, keepdims=True)
        xyz.append(cur_xyz[par] + rng.normal(0, 1, (k * max_child, 3)) * sc)
        logs.append(cur_log[par] - math.log(2.0) + rng.normal(0, 0.1, (k * max_child, 3)))
        node_index = np.concatenate([node_index, np.full(k * max_child, -1, np.int64)])
        depth = np.concatenate([depth, np.repeat(depth[pick] + 1, max_child)])
        parent = np.concatenate([parent, par])
        N += k * max_child
    tree = np.array(rows, np.int64).reshape(-1, max_child) if rows else np.zeros((0, max_child), np.int64)
    # removals: leaves that are not roots lose their slot
    if tree.size:
        cand = np.nonzero((node_index == -1) & (parent >= 0))[0]
        gone = cand[rng.random(cand.size) < remove_frac]
        for g in gone:
            row = tree[node_index[parent[g]]]
            row[row == g] = -1
        empty = np.nonzero((tree == -1).all(1))[0]
        owner = np.full(tree.shape[0], -1, np.int64)
        owner[node_index[node_index >= 0]] = np.nonzero(node_index >= 0)[0]
        node_index[owner[empty]] = -1
    q = rng.normal(0, 1, (N, 4))
    q[rng.random(N) < 0.02] = 0.0          # zero quaternions: F.normalize gives 0
    return dict(xyz=np.concatenate(xyz).astype(np.float32), scaling=np.concatenate(logs).astype(np.float32),
                rotation=q.astype(np.float32), node_index=node_index.astype(np.int32),
                tree=tree.astype(np.int32), depth=depth.astype(np.int32),
                index_parent=parent.astype(np.int32))


def chain_tree(length=36, seed=0, dist=10.0):
    """One root and a single-child chain of `length` nodes (depth length-1),
    all large on screen: the cut runs into the level limit."""
    rng = np.random.default_rng(seed)
    N = length
    node_index = np.arange(N, dtype=np.int32)
    node_index[-1] = -1
    tree = np.arange(1, N, dtype=np.int32).reshape(-1, 1)
    xyz = np.zeros((N, 3), np.float32)
    xyz[:, 1] = dist
    xyz[:, 0] = rng.normal(0, 0.1, N)
    scaling = np.full((N, 3), math.log(2.0), np.float32)
    rotation = np.tile(np.array([1, 0, 0, 0], np.float32), (N, 1))
    return dict(xyz=xyz, scaling=scaling, rotation=rotation, node_index=node_index, tree=tree,
                depth=np.arange(N, dtype=np.int32), index_parent=np.arange(-1, N - 1, dtype=np.int32))
