"""TensorTree.traverse of LoG (LoG/model/tensor_tree.py:133-188) on the GPU.

The reference walks the tree level by level in Python: per level it gathers
the frontier's children, calls Gaussian.compute_radius on them
(LoG/model/level_of_gaussian.py:66-88: exp of the raw scales, F.normalize of
the raw rotations, then compute_radius_cuda), splits them into kept and
continuing nodes and waits for the host (`flag_go_to_next.sum() == 0`) before
the next level.  Here each level is three kernel launches that read the
level's size from device memory (the radius and activation fused into the
decision), and the host waits once, for the size of the result.

`traverse(tree, model, root_index, camera, max_depth)` takes the reference's
arguments: `tree` with the TensorTree buffers (`node_index`, `tree`, `depth`,
`min_resolution_pixel`), `model` a LoG Gaussian (`xyz`, `scaling`,
`rotation`, exp scaling activation), `camera` anything with LoG's
`raster_settings` (viewmatrix, projmatrix, tanfovx, tanfovy, image_width,
image_height).  `DeviceTraverse` is a mixin that gives a TensorTree subclass
this method.  The ids come back as int64 in the reference's order.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import call, ptr

_DEPTH_CACHE_ATTR = "_fgs_depth_bound"


def tree_traverse(xyz, scaling, rotation, node_index, tree, root_index, viewmatrix, projmatrix,
                  focal_x, focal_y, tanfovx, tanfovy, min_resolution_pixel=3.0, max_depth=1000,
                  level_hint=None):
    """The cut of TensorTree.traverse from raw tensors.

    xyz [N,3], scaling [N,3] raw (log) scales, rotation [N,4] raw quaternions
    (float32, on the GPU); node_index [N] (row in `tree`, -1 for a leaf);
    tree [n_nodes, C] child ids (-1 for an empty slot); root_index [R] ids.
    level_hint bounds the number of levels run (the tree's depth below the
    roots or more); None runs min(30, max_depth) levels.  Returns int64 [n]."""
    dev = xyz.device
    _lib.require_gpu(xyz, scaling, rotation, node_index, tree, root_index)
    if xyz.dim() != 2 or xyz.shape[1] != 3 or scaling.shape != xyz.shape or rotation.shape != (xyz.shape[0], 4):
        raise ValueError("xyz and scaling must be [N, 3], rotation [N, 4]")
    if tree.dim() != 2 or tree.shape[1] < 1:
        raise ValueError("tree must be [n_nodes, max_child]")
    N = int(xyz.shape[0])
    if node_index.numel() != N:
        raise ValueError("node_index must have one entry per point")
    C = int(tree.shape[1])
    f32 = lambda t: t.to(device=dev, dtype=torch.float32).contiguous()
    i32 = lambda t: t.to(device=dev, dtype=torch.int32).contiguous()
    roots = root_index.to(device=dev, dtype=torch.int64).contiguous().view(-1)
    R = int(roots.numel())
    lib = _lib.load()
    ws = torch.empty((max(1, lib.fgs_tree_traverse_ws_bytes(N, C, R)),), dtype=torch.uint8, device=dev)
    xyz_c, sc_c, rot_c = f32(xyz), f32(scaling), f32(rotation)
    ni_c, tree_c = i32(node_index), i32(tree)
    vm, pm = f32(viewmatrix).view(-1), f32(projmatrix).view(-1)
    count = ctypes.c_int64(0)
    stream = _lib.stream_handle(dev)
    hint = -1 if level_hint is None else max(0, int(level_hint))
    call("fgs_tree_traverse", N, ptr(xyz_c), ptr(sc_c), ptr(rot_c), ptr(ni_c), ptr(tree_c),
         int(tree.shape[0]), C, ptr(roots), R, ptr(pm), ptr(vm), float(focal_x), float(focal_y),
         float(tanfovx), float(tanfovy), float(min_resolution_pixel),
         int(max(-1, min(int(max_depth), 1 << 30))), hint, ws.data_ptr(), ctypes.addressof(count),
         stream)
    out = torch.empty((int(count.value),), dtype=torch.int64, device=dev)
    if out.numel():
        call("fgs_tree_traverse_result", N, C, R, ws.data_ptr(), int(count.value), out.data_ptr(), stream)
    return out


def _depth_bound(tree):
    """max(depth) of the tree, cached on the tree object until its depth
    buffer changes (TensorTree.split / remove replace it with set_)."""
    depth = tree.depth
    key = (depth.data_ptr(), depth.numel(), depth._version)
    cached = getattr(tree, _DEPTH_CACHE_ATTR, None)
    if cached is not None and cached[0] == key:
        return cached[1]
    bound = int(depth.max().item()) if depth.numel() else 0
    try:
        setattr(tree, _DEPTH_CACHE_ATTR, (key, bound))
    except AttributeError:
        pass
    return bound


def traverse(tree, model, root_index, camera, max_depth=1000):
    """TensorTree.traverse(model, root_index, camera, max_depth)
    (LoG/model/tensor_tree.py:166-188) on the GPU: same arguments, same ids
    in the same order (int64)."""
    act = getattr(model, "activation", None)
    if act is not None and getattr(act, "scaling_activation", torch.exp) is not torch.exp:
        raise NotImplementedError("traverse: only the exp scaling activation is fused")
    rs = camera.raster_settings
    W, H = int(rs.image_width), int(rs.image_height)
    fx = W / (2.0 * rs.tanfovx)
    fy = H / (2.0 * rs.tanfovy)
    nodes = tree.tree
    if nodes.numel() == 0:
        nodes = torch.full((1, max(1, int(getattr(tree, "max_child", 1)))), -1, dtype=torch.int32,
                           device=model.xyz.device)
    with torch.no_grad():
        return tree_traverse(model.xyz.detach(), model.scaling.detach(), model.rotation.detach(),
                             tree.node_index, nodes, root_index, rs.viewmatrix, rs.projmatrix, fx, fy,
                             rs.tanfovx, rs.tanfovy, float(tree.min_resolution_pixel), max_depth,
                             level_hint=_depth_bound(tree))


class DeviceTraverse:
    """Mixin for a TensorTree subclass: `traverse` runs on the GPU.

        class Tree(DeviceTraverse, TensorTree): pass
    """

    def traverse(self, model, root_index, camera, max_depth=1000):
        return traverse(self, model, root_index, camera, max_depth)
