"""ctypes binding of libfiltergs_amd.so (the gfx950 C ABI, include/filtergs_amd.h).

torch supplies device memory and the current HIP stream; every pointer handed
to the library is ``tensor.data_ptr()`` of a contiguous float32/int tensor on
the GPU.  There is no CPU fallback: when the library (or a GPU) is missing the
calls raise.
"""
from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfiltergs_amd.so")
_lib = None

c_void_p = ctypes.c_void_p
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_float = ctypes.c_float
c_size_t = ctypes.c_size_t

FGS_VARIANT_WODILATE = 0
FGS_VARIANT_DILATE = 1
FGS_GRAD_REFERENCE = 0
FGS_GRAD_EXACT = 1


class RasterSettings(ctypes.Structure):
    """fgs_raster_settings."""
    _fields_ = [
        ("image_height", c_int32),
        ("image_width", c_int32),
        ("tanfovx", c_float),
        ("tanfovy", c_float),
        ("bg", c_void_p),
        ("scale_modifier", c_float),
        ("viewmatrix", c_void_p),
        ("projmatrix", c_void_p),
        ("sh_degree", c_int32),
        ("campos", c_void_p),
        ("prefiltered", c_int32),
        ("debug", c_int32),
        ("variant", c_int32),
        ("use_filter", c_int32),
        ("ogfs_epsilon_scale", c_float),
        ("ogfs_epsilon_max", c_float),
        ("ogfs_energy_floor", c_float),
        ("ogfs_radius_ratio_min", c_float),
        ("ogfs_enable_stats", c_int32),
        ("grad_mode", c_int32),
    ]


class Gaussians(ctypes.Structure):
    """fgs_gaussians."""
    _fields_ = [
        ("P", c_int32),
        ("M", c_int32),
        ("means3D", c_void_p),
        ("shs", c_void_p),
        ("colors_precomp", c_void_p),
        ("opacities", c_void_p),
        ("scales", c_void_p),
        ("rotations", c_void_p),
        ("cov3D_precomp", c_void_p),
    ]


class FilterGSRows(ctypes.Structure):
    """fgs_filtergs_rows."""
    _fields_ = [
        ("N", c_int32),
        ("index", c_void_p),
        ("n_src", c_int64),
        ("xyz", c_void_p),
        ("scaling", c_void_p),
        ("rotation", c_void_p),
        ("opacity", c_void_p),
        ("colors", c_void_p),
        ("shs", c_void_p),
        ("rest_coeffs", c_int32),
        ("sh_degree", c_int32),
        ("campos", c_void_p),
    ]


class BatchView(ctypes.Structure):
    """fgs_batch_view."""
    _fields_ = [
        ("image_height", c_int32),
        ("image_width", c_int32),
        ("tanfovx", c_float),
        ("tanfovy", c_float),
        ("viewmatrix", c_float * 16),
        ("projmatrix", c_float * 16),
        ("campos", c_float * 3),
        ("bg", c_float * 3),
    ]


class BatchSizes(ctypes.Structure):
    """fgs_batch_sizes."""
    _fields_ = [
        ("views", c_int32),
        ("P", c_int32),
        ("tiles", c_int64),
        ("max_view_tiles", c_int64),
        ("pixels", c_int64),
        ("plan_bytes", c_size_t),
        ("geom_bytes", c_size_t),
        ("image_bytes", c_size_t),
        ("backward_ws_bytes", c_size_t),
    ]


_SIGS = {
    "fgs_geom_bytes": (c_size_t, [c_int32]),
    "fgs_image_bytes": (c_size_t, [c_int32, c_int32]),
    "fgs_binning_bytes": (c_size_t, [c_int64, c_int32, c_int32]),
    "fgs_forward_geometry": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fgs_forward_render": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p]),
    "fgs_forward_render_stats": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                           c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           c_void_p, c_void_p, ctypes.c_float, c_void_p]),
    "fgs_stats_words": (c_size_t, [c_int32, c_int32]),
    "fgs_sparse_adam": (c_int32, [c_int64, c_int32, c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                  c_void_p, c_void_p, c_void_p, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_double, c_void_p, c_void_p]),
    "fgs_counter_update": (c_int32, [c_int64, c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                     c_void_p, c_int64] + [c_void_p] * 10),
    "fgs_compute_visibility_radius_parent": (c_int32, [c_int32] + [c_void_p] * 7 + [c_float] * 7 +
                                             [c_void_p] * 5),
    "fgs_rebuild_ancestor_path": (c_int32, [c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                            c_void_p]),
    "fgs_activate_backward": (c_int32, [c_int64] + [c_void_p] * 4 + [c_int32, c_int32] + [c_void_p] * 11),
    "fgs_tree_traverse_ws_bytes": (c_size_t, [c_int32, c_int32, c_int32]),
    "fgs_tree_traverse": (c_int32, [c_int32] + [c_void_p] * 5 + [c_int32, c_int32, c_void_p, c_int32,
                                    c_void_p, c_void_p] + [c_float] * 5 + [c_int32, c_int32,
                                    c_void_p, c_void_p, c_void_p]),
    "fgs_tree_traverse_result": (c_int32, [c_int32, c_int32, c_int32, c_void_p, c_int64, c_void_p,
                                           c_void_p]),
    "fgs_filtergs_geometry": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fgs_filtergs_render": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                      ctypes.c_float, c_void_p]),
    "fgs_read_sorted": (c_int32, [c_void_p, c_int32, c_int64, c_void_p, c_void_p, c_void_p,
                                  c_void_p]),
    "fgs_image_ranges": (c_void_p, [c_void_p, c_int32, c_int32]),
    "fgs_geom_offsets": (None, [c_int32, c_void_p]),
    "fgs_image_offsets": (None, [c_int32, c_int32, c_void_p]),
    "fgs_backward_ws_bytes": (c_size_t, [c_int32]),
    "fgs_backward": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fgs_backward_noalloc": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_int64, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fgs_mark_visible": (c_int32, [c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fgs_compute_radius": (c_int32, [c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_float, c_float, c_int32, c_int32, c_void_p, c_void_p]),
    "fgs_compute_radius_v2": (c_int32, [c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_float, c_float, c_float, c_float, c_void_p,
                                        c_void_p]),
    "fgs_lod_ws_bytes": (c_size_t, [c_int32]),
    "fgs_compute_visibility_radius": (c_int32, [c_int32, c_void_p, c_void_p, c_void_p, c_void_p,
                                                c_void_p, c_void_p, c_void_p, c_int32, c_float,
                                                c_float, c_float, c_float, c_float, c_float,
                                                c_float, c_void_p, c_void_p, c_void_p, c_void_p,
                                                c_void_p]),
    "fgs_visible_flag": (c_int32, [c_int32, c_int32, c_void_p, c_void_p, c_float, c_void_p,
                                   c_void_p, c_void_p, c_void_p]),
    "fgs_gather_attributes": (c_int32, [c_int64, c_void_p, c_int64, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
                                        c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_void_p]),
    "fgs_activate_gather": (c_int32, [c_int64, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_void_p,
                                      c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p]),
    "fgs_batch_sizes_query": (c_int32, [c_void_p, c_int32, c_int32, c_void_p]),
    "fgs_batch_binning_bytes": (c_size_t, [c_int64]),
    "fgs_batch_offsets": (c_int32, [c_void_p, c_int64, c_void_p]),
    "fgs_batch_plan": (c_int32, [c_void_p, c_int32, c_void_p, c_void_p]),
    "fgs_batch_forward": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "fgs_batch_backward": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_int64, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p, c_void_p]),
    "fgs_profile_enable": (c_int32, [c_int32]),
    "fgs_profile_read": (c_int32, [c_void_p, c_void_p, c_int32, c_int32]),
    "fgs_set_work_shape": (c_int32, [c_int32, c_int32]),
    "fgs_get_work_shape": (c_int32, [c_void_p]),
    "fgs_set_blend_walk": (c_int32, [c_int32]),
    "fgs_set_backward_layout": (c_int32, [c_int32]),
    "fgs_get_backward_layout": (c_int32, []),
    "fgs_set_forward_layout": (c_int32, [c_int32]),
    "fgs_set_xcd_chunk": (c_int32, [c_int32]),
    "fgs_set_tile_queue": (c_int32, [c_int32, c_int32]),
    "fgs_get_tile_queue": (c_int32, []),
    "fgs_get_forward_layout": (c_int32, []),
    "fgs_last_error": (ctypes.c_char_p, []),
    "fgs_abi_version": (c_int32, []),
    "fgs_ogfs_log_lambda": (c_float, [c_float, c_float]),
}

EXPORTED = tuple(_SIGS)


class FGSError(RuntimeError):
    pass


def load():
    """Load the shared library (no GPU needed) and bind every entry point."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FGSError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' "
            "or make -C filtergs_amd/csrc)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def call(name, *args):
    rc = getattr(load(), name)(*args)
    if rc != 0:
        msg = load().fgs_last_error().decode(errors="replace")
        raise FGSError(f"{name} failed (code {rc}): {msg}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL).  The tensor must be contiguous."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor passed to the C ABI must be contiguous")
    return t.data_ptr() if t.numel() > 0 else None


def stream_handle(device=None):
    return torch.cuda.current_stream(device).cuda_stream


def require_gpu(*tensors):
    if not torch.cuda.is_available():
        raise FGSError("filtergs_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    for t in tensors:
        if t is not None and torch.is_tensor(t) and t.numel() > 0 and not t.is_cuda:
            raise ValueError("all tensors must live on the GPU")


STAGES = ("preprocess", "depth_sort", "tile_offsets", "duplicate", "tile_sort", "ranges",
          "blend_fwd", "blend_bwd", "preprocess_bwd", "blend_fwd_kernel", "blend_bwd_kernel")


def profile_enable(on=True):
    call("fgs_profile_enable", int(bool(on)))


def profile_read(reset=True):
    """{stage: (total_ms, calls)} accumulated since the last reset."""
    n = len(STAGES)
    ms = (ctypes.c_double * n)()
    calls = (ctypes.c_uint64 * n)()
    load().fgs_profile_read(ms, calls, n, int(bool(reset)))
    return {STAGES[k]: (float(ms[k]), int(calls[k])) for k in range(n)}


def set_work_shape(fwd_batch=256, bwd_batch=128):
    """Blend work shape: Gaussians per LDS batch (forward 128/256, backward 64/128/256)."""
    call("fgs_set_work_shape", int(fwd_batch), int(bwd_batch))


def set_backward_layout(layout=1):
    """Backward tile layout: 0 four waves per tile (one strip each), 1 one
    wave per tile with four pixels per lane (fgs_set_backward_layout)."""
    call("fgs_set_backward_layout", int(layout))


def get_backward_layout():
    return int(load().fgs_get_backward_layout())


def set_forward_layout(layout=2):
    """Forward tile layout (without statistics): 0 four waves per tile, 1 one
    wave per tile with four pixels per lane, 2 (default) layout 1 for
    launches of at least 12288 tiles (fgs_set_forward_layout)."""
    call("fgs_set_forward_layout", int(layout))


def get_forward_layout():
    return int(load().fgs_get_forward_layout())


def set_xcd_chunk(chunk=32):
    """Tile order of the batched one-wave blend kernels: runs of `chunk`
    consecutive tiles per XCD (0: launch order) (fgs_set_xcd_chunk)."""
    call("fgs_set_xcd_chunk", int(chunk))


def set_tile_queue(chunk=0, by_cost=True):
    """Batched one-wave blends as persistent launches over a (view, tile)
    queue of runs of `chunk` consecutive tiles, dealt to one queue per XCD;
    `by_cost` orders the runs heaviest first.  0 launches one workgroup per
    tile (fgs_set_tile_queue)."""
    call("fgs_set_tile_queue", int(chunk), int(bool(by_cost)))


def get_tile_queue():
    return int(load().fgs_get_tile_queue())


def set_blend_walk(mode=0):
    """Forward blend walk: 0 adaptive (default), 1 wave-uniform, 2 divergent
    (same images, different speed)."""
    call("fgs_set_blend_walk", int(mode))


def get_work_shape():
    out = (ctypes.c_int32 * 2)()
    call("fgs_get_work_shape", out)
    return tuple(int(v) for v in out)


def host_copy(args):
    """Every tensor of ``args`` copied to host memory (other items as they
    are): the arguments of a call, taken before the call can change them."""
    return tuple(a.detach().cpu().clone() if isinstance(a, torch.Tensor) else a for a in args)


def guarded_call(fn, args, kwargs, debug, dump_path, stage):
    """``fn(*args, **kwargs)``.  With ``debug`` set, a host copy of ``args`` is
    taken first; if the call raises, the copy is written to ``dump_path`` with
    torch.save and the exception propagates.  This is the plugin surfaces'
    debug convention: the snapshot files are named snapshot_fw.dump and
    snapshot_bw.dump (antialias/diff_gaussian_rasterization*/__init__.py)."""
    if not debug:
        return fn(*args, **kwargs)
    saved = host_copy(args)
    try:
        return fn(*args, **kwargs)
    except Exception:
        torch.save(saved, dump_path)
        print(f"\n[filtergs_amd] the rasterizer {stage} raised; its arguments are saved in {dump_path}")
        raise
