"""The FilterGS/cuda surface: LoD cut, 3-sigma radius, NDC frustum flags and
attribute gather, with the signatures, outputs and errors of
FilterGS/cuda/compute_radius.py:23-45, FilterGS/cuda/visible_flag.py:19-207
and FilterGS/cuda/gather_attributes.py:20-40.
"""
from __future__ import annotations

import time
from types import SimpleNamespace

import torch

from . import _lib
from ._lib import call, ptr


def _f32c(t, dev):
    return t.to(device=dev, dtype=torch.float32).contiguous()


def _compute_radius(means3D, scales, rotations, projmatrix, viewmatrix, focal_x, focal_y,
                    tan_fovx, tan_fovy):
    """compute_radius (compute_radius_kernel_v2.cu:355-380): scales and quaternions used as
    given (LoG passes exp / normalize, level_of_gaussian.py:66-84; the FilterGS fallback,
    filtergs.py:261-268, passes its raw buffers)."""
    P = int(means3D.size(0))
    radii = torch.zeros((P,), dtype=means3D.dtype, device=means3D.device)
    if P == 0:
        return radii
    _lib.require_gpu(means3D)
    dev = means3D.device
    m, s, r = _f32c(means3D, dev), _f32c(scales, dev), _f32c(rotations, dev)
    pm, vm = _f32c(projmatrix, dev), _f32c(viewmatrix, dev)
    call("fgs_compute_radius_v2", P, ptr(m), ptr(s), ptr(r), ptr(pm), ptr(vm), float(focal_x),
         float(focal_y), float(tan_fovx), float(tan_fovy), radii.data_ptr(), _lib.stream_handle(dev))
    return radii


def _compute_visibility_radius_ext(means3D, scales, rotations, projmatrix, viewmatrix,
                                   tree_node_index, ancestor_path, focal_x, focal_y, tan_fovx,
                                   tan_fovy, padding, radius_max, radius_filter_threshold):
    """compute_visibility_radius host function (compute_radius_kernel_v2.cu:382-430)."""
    P = int(means3D.size(0))
    dev = means3D.device
    mask = torch.zeros((P,), dtype=torch.bool, device=dev)
    radius2d = torch.full((P,), -1.0, dtype=torch.float32, device=dev)
    depth = torch.zeros((P,), dtype=torch.float32, device=dev)
    if P == 0:
        return mask, radius2d, depth
    _lib.require_gpu(means3D)
    L = int(ancestor_path.size(1)) if ancestor_path.dim() == 2 else 0
    m, s, r = _f32c(means3D, dev), _f32c(scales, dev), _f32c(rotations, dev)
    pm, vm = _f32c(projmatrix, dev), _f32c(viewmatrix, dev)
    ni = tree_node_index.to(device=dev, dtype=torch.int32).contiguous()
    ap = ancestor_path.to(device=dev, dtype=torch.int32).contiguous()
    ws = torch.empty((_lib.load().fgs_lod_ws_bytes(P),), dtype=torch.uint8, device=dev)
    call("fgs_compute_visibility_radius", P, ptr(m), ptr(s), ptr(r), ptr(pm), ptr(vm), ptr(ni),
         ptr(ap), L, float(focal_x), float(focal_y), float(tan_fovx), float(tan_fovy),
         float(padding), float(radius_max), float(radius_filter_threshold), mask.data_ptr(),
         radius2d.data_ptr(), depth.data_ptr(), ws.data_ptr(), _lib.stream_handle(dev))
    return mask, radius2d, depth


# The object FilterGS imports as ``compute_radius_module`` (compute_radius.py:7-20).
compute_radius_module = SimpleNamespace(compute_radius=_compute_radius,
                                        compute_visibility_radius=_compute_visibility_radius_ext)


def compute_visibility_radius(means3d, scales, rotations, projmatrix, viewmatrix, tree_node_index,
                              ancestor_path, focal_x, focal_y, tan_fovx, tan_fovy, radius_max,
                              padding=0.05, radius_filter_threshold=100.0):
    """Compute visibility mask, 2D radius, and depth with optional LOD filtering."""
    tensors = (means3d, scales, rotations, projmatrix, viewmatrix, tree_node_index, ancestor_path)
    if not all(t.is_cuda for t in tensors):
        raise ValueError("All input tensors must be CUDA tensors")
    return compute_radius_module.compute_visibility_radius(
        means3d, scales, rotations, projmatrix, viewmatrix, tree_node_index, ancestor_path,
        float(focal_x), float(focal_y), float(tan_fovx), float(tan_fovy), float(padding),
        float(radius_max), float(radius_filter_threshold))


def compute_visibility_radius_parent(means3d, scales, rotations, projmatrix, viewmatrix,
                                     tree_node_index, index_parent, focal_x, focal_y, tan_fovx,
                                     tan_fovy, radius_max, padding=0.05, radius_filter_threshold=100.0):
    """compute_visibility_radius with the tree's parent links (tree.index_parent,
    -1 at a root) instead of the [N, L] ancestor rows: the same mask as the
    rows rebuild_ancestor_path makes from those links, reading 4 bytes per
    node instead of 4 L."""
    tensors = (means3d, scales, rotations, projmatrix, viewmatrix, tree_node_index, index_parent)
    if not all(t.is_cuda for t in tensors):
        raise ValueError("All input tensors must be CUDA tensors")
    P = int(means3d.size(0))
    dev = means3d.device
    mask = torch.zeros((P,), dtype=torch.bool, device=dev)
    radius2d = torch.full((P,), -1.0, dtype=torch.float32, device=dev)
    depth = torch.zeros((P,), dtype=torch.float32, device=dev)
    if P == 0:
        return mask, radius2d, depth
    m, s, r = _f32c(means3d, dev), _f32c(scales, dev), _f32c(rotations, dev)
    pm, vm = _f32c(projmatrix, dev), _f32c(viewmatrix, dev)
    ni = tree_node_index.to(device=dev, dtype=torch.int32).contiguous()
    ip = index_parent.to(device=dev, dtype=torch.int32).contiguous()
    ws = torch.empty((_lib.load().fgs_lod_ws_bytes(P),), dtype=torch.uint8, device=dev)
    call("fgs_compute_visibility_radius_parent", P, ptr(m), ptr(s), ptr(r), ptr(pm), ptr(vm), ptr(ni),
         ptr(ip), float(focal_x), float(focal_y), float(tan_fovx), float(tan_fovy), float(padding),
         float(radius_max), float(radius_filter_threshold), mask.data_ptr(), radius2d.data_ptr(),
         depth.data_ptr(), ws.data_ptr(), _lib.stream_handle(dev))
    return mask, radius2d, depth


def rebuild_ancestor_path(state_dict, device=None):
    """rebuild_ancestor_path of apps/ancestor.py:21-81 with the rows built on
    the GPU: reads tree.index_parent and tree.depth, writes
    state_dict["tree.ancestor_path"] (and "tree.ancestor" when present) with
    the width of an existing ancestor buffer, else depth.max() + 1; returns
    (ancestor_path, max_depth).  The rows land on the device of
    tree.index_parent unless ``device`` names a GPU to build them on, in
    which case they stay there."""
    for key in ("tree.index_parent", "tree.depth"):
        if key not in state_dict:
            raise KeyError(f"Missing '{key}' in checkpoint state_dict")
    index_parent = state_dict["tree.index_parent"]
    depth = state_dict["tree.depth"]
    home = index_parent.device
    dev = torch.device(device) if device is not None else (home if home.type == "cuda" else torch.device("cuda"))
    _lib.require_gpu(torch.empty(0, device=dev))
    ancestor_key = next((k for k in ("tree.ancestor_path", "tree.ancestor") if k in state_dict), None)
    N = int(depth.numel())
    if ancestor_key is not None and state_dict[ancestor_key].dim() == 2:
        L = int(state_dict[ancestor_key].shape[1])
        out_dtype = state_dict[ancestor_key].dtype
    else:
        L = int(depth.max().item()) + 1 if N > 0 else 1
        out_dtype = torch.int32
    ip = index_parent.to(device=dev, dtype=torch.int32).contiguous()
    dp = depth.to(device=dev, dtype=torch.int32).contiguous()
    out = torch.empty((N, L), dtype=torch.int32, device=dev)
    bad = torch.zeros((1,), dtype=torch.int32, device=dev)
    max_depth = 0
    if N > 0 and L > 0:
        call("fgs_rebuild_ancestor_path", N, L, ptr(ip), ptr(dp), out.data_ptr(), bad.data_ptr(),
             _lib.stream_handle(dev))
        if int(bad.item()):
            raise IndexError(f"a node's depth exceeds the {L} ancestor columns")
        max_depth = int(dp.max().item())
    else:
        out.fill_(-1)
    target = dev if device is not None else home
    out = out.to(device=target, dtype=out_dtype)
    state_dict["tree.ancestor_path"] = out
    if "tree.ancestor" in state_dict:
        state_dict["tree.ancestor"] = out
    return out, max_depth


def _visible_flag(xyz, proj_matrix, padding, batch):
    dev = xyz.device
    B = int(xyz.size(0)) if batch else 1
    N = int(xyz.size(1)) if batch else int(xyz.size(0))
    shape = (B, N) if batch else (N,)
    valid = torch.zeros(shape, dtype=torch.bool, device=dev)
    depth = torch.zeros(shape, dtype=torch.float32, device=dev)
    p_proj = torch.zeros(shape + (3,), dtype=torch.float32, device=dev)
    if B * N == 0:
        return valid, depth, p_proj
    x = _f32c(xyz, dev)
    pm = _f32c(proj_matrix, dev)
    call("fgs_visible_flag", B, N, ptr(x), ptr(pm), float(padding), valid.data_ptr(),
         depth.data_ptr(), p_proj.data_ptr(), _lib.stream_handle(dev))
    return valid, depth, p_proj


def visible_flag_cuda(xyz, camera, padding=0.05, use_optimized=True):
    """(valid_flag, depth, p_proj) for xyz [N,3]; camera dict with 'full_proj_transform'."""
    if not isinstance(xyz, torch.Tensor):
        raise ValueError("xyz must be a torch.Tensor")
    if not xyz.is_cuda:
        raise ValueError("xyz must be on GPU")
    if xyz.dim() != 2 or xyz.size(1) != 3:
        raise ValueError("xyz must have shape [N, 3]")
    if not isinstance(camera, dict) or 'full_proj_transform' not in camera:
        raise ValueError("camera must be a dict containing 'full_proj_transform' key")
    proj_matrix = camera['full_proj_transform']
    if not proj_matrix.is_cuda:
        proj_matrix = proj_matrix.cuda()
    if proj_matrix.dim() != 2 or proj_matrix.size(0) != 4 or proj_matrix.size(1) != 4:
        raise ValueError("full_proj_transform must have shape [4, 4]")
    try:
        return _visible_flag(xyz, proj_matrix, padding, batch=False)
    except Exception as e:
        raise RuntimeError(f"CUDA computation failed: {str(e)}")


def visible_flag_cuda_direct(xyz, proj_matrix, padding=0.05, use_optimized=True):
    if not isinstance(xyz, torch.Tensor):
        raise ValueError("xyz must be a torch.Tensor")
    if not xyz.is_cuda:
        raise ValueError("xyz must be on GPU")
    if xyz.dim() != 2 or xyz.size(1) != 3:
        raise ValueError("xyz must have shape [N, 3]")
    if not isinstance(proj_matrix, torch.Tensor):
        raise ValueError("proj_matrix must be a torch.Tensor")
    if not proj_matrix.is_cuda:
        raise ValueError("proj_matrix must be on GPU")
    if proj_matrix.dim() != 2 or proj_matrix.size(0) != 4 or proj_matrix.size(1) != 4:
        raise ValueError("proj_matrix must have shape [4, 4]")
    try:
        return _visible_flag(xyz, proj_matrix, padding, batch=False)
    except Exception as e:
        raise RuntimeError(f"CUDA computation failed: {str(e)}")


def visible_flag_cuda_batch(xyz, camera, padding=0.05, use_optimized=True):
    """Batch version for xyz [B, N, 3] with one shared projection matrix."""
    if xyz.dim() != 3:
        raise ValueError("xyz must have shape [B, N, 3] for batch processing")
    proj_matrix = camera['full_proj_transform']
    if not proj_matrix.is_cuda:
        proj_matrix = proj_matrix.cuda()
    if proj_matrix.dim() != 2 or proj_matrix.size(0) != 4 or proj_matrix.size(1) != 4:
        raise ValueError("full_proj_transform must have shape [4, 4]")
    try:
        return _visible_flag(xyz, proj_matrix, padding, batch=True)
    except Exception as e:
        raise RuntimeError(f"CUDA batch computation failed: {str(e)}")


def visible_flag_cuda_batch_direct(xyz, proj_matrix, padding=0.05, use_optimized=True):
    if xyz.dim() != 3:
        raise ValueError("xyz must have shape [B, N, 3] for batch processing")
    if not isinstance(proj_matrix, torch.Tensor):
        raise ValueError("proj_matrix must be a torch.Tensor")
    if not proj_matrix.is_cuda:
        raise ValueError("proj_matrix must be on GPU")
    if proj_matrix.dim() != 2 or proj_matrix.size(0) != 4 or proj_matrix.size(1) != 4:
        raise ValueError("proj_matrix must have shape [4, 4]")
    try:
        return _visible_flag(xyz, proj_matrix, padding, batch=True)
    except Exception as e:
        raise RuntimeError(f"CUDA batch computation failed: {str(e)}")


def benchmark_visible_flag(xyz, camera, padding=0.05, num_runs=100):
    """Launch-loop timing (visible_flag_kernel.cu:290-356): (total_ms, avg_ms, runs, N)."""
    proj = camera['full_proj_transform']
    for _ in range(5):
        visible_flag_cuda(xyz, camera, padding)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(num_runs):
        visible_flag_cuda(xyz, camera, padding)
    torch.cuda.synchronize()
    total = (time.perf_counter() - t0) * 1000.0
    del proj
    return torch.tensor([total, total / num_runs, float(num_runs), float(xyz.size(0))],
                        dtype=torch.float32, device=xyz.device)


def gather_attributes(index, xyz, scaling, rotation, opacity, colors, shs=None):
    """Row gather of the visible Gaussians' attributes -> dict."""
    tensors = [index, xyz, scaling, rotation, opacity, colors]
    if shs is not None:
        tensors.append(shs)
    for tensor in tensors:
        if tensor is not None:
            assert tensor.is_cuda, "All input tensors must be CUDA tensors"
    if index.dtype != torch.long:
        index = index.to(dtype=torch.long)
    index = index.contiguous()
    dev = xyz.device
    N = int(index.numel())
    n_src = int(xyz.size(0))
    src = [_f32c(t, dev) for t in (xyz, scaling, rotation, opacity, colors)]
    outs = [torch.empty((N,) + tuple(t.shape[1:]), dtype=torch.float32, device=dev) for t in src]
    sh_src = _f32c(shs, dev) if shs is not None else None
    sh_out = (torch.empty((N,) + tuple(shs.shape[1:]), dtype=torch.float32, device=dev)
              if shs is not None else None)
    sh_floats = int(sh_src[0].numel()) if (sh_src is not None and sh_src.size(0) > 0) else 0
    bad = torch.zeros((1,), dtype=torch.int32, device=dev)
    if N > 0:
        call("fgs_gather_attributes", N, index.data_ptr(), n_src, *[ptr(t) for t in src],
             ptr(sh_src), sh_floats, *[ptr(t) for t in outs], ptr(sh_out), bad.data_ptr(),
             _lib.stream_handle(dev))
        if int(bad.item()):
            raise RuntimeError("gather_attributes: index out of range")
    keys = ['xyz', 'scaling', 'rotation', 'opacity', 'colors']
    result = {key: outs[i] for i, key in enumerate(keys)}
    if shs is not None:
        result['shs'] = sh_out
    return result


__all__ = ['compute_radius_module', 'compute_visibility_radius', 'visible_flag_cuda',
           'visible_flag_cuda_direct', 'visible_flag_cuda_batch', 'visible_flag_cuda_batch_direct',
           'benchmark_visible_flag', 'gather_attributes']
