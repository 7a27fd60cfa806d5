"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module; the product package filtergs_amd never imports it.  All arrays are
numpy; the f32 functions take float32 arrays, the f64 ones float64 arrays.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfgs_oracle.so")
_lib = None


class OrcSettings(ctypes.Structure):
    _fields_ = [
        ("image_height", ctypes.c_int32),
        ("image_width", ctypes.c_int32),
        ("tanfovx", ctypes.c_float),
        ("tanfovy", ctypes.c_float),
        ("bg", ctypes.c_float * 3),
        ("scale_modifier", ctypes.c_float),
        ("viewmatrix", ctypes.c_float * 16),
        ("projmatrix", ctypes.c_float * 16),
        ("campos", ctypes.c_float * 3),
        ("sh_degree", ctypes.c_int32),
        ("prefiltered", ctypes.c_int32),
        ("variant", ctypes.c_int32),
        ("use_filter", ctypes.c_int32),
        ("ogfs_epsilon_scale", ctypes.c_float),
        ("ogfs_epsilon_max", ctypes.c_float),
        ("ogfs_energy_floor", ctypes.c_float),
        ("ogfs_radius_ratio_min", ctypes.c_float),
        ("grad_mode", ctypes.c_int32),
    ]


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the .so path."""
    if force or not os.path.exists(_LIB_PATH):
        cmd = ["make", "-C", _HERE]
        if force:
            cmd.insert(1, "-B")
        subprocess.run(cmd, check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def settings(image_height, image_width, tanfovx, tanfovy, viewmatrix, projmatrix, campos,
             bg=(0.0, 0.0, 0.0), scale_modifier=1.0, sh_degree=0, variant=0, use_filter=True,
             prefiltered=False, ogfs_epsilon_scale=0.2, ogfs_epsilon_max=0.2,
             ogfs_energy_floor=0.05, ogfs_radius_ratio_min=0.7, grad_mode=0) -> OrcSettings:
    s = OrcSettings()
    s.image_height = int(image_height)
    s.image_width = int(image_width)
    s.tanfovx = float(tanfovx)
    s.tanfovy = float(tanfovy)
    s.bg = (ctypes.c_float * 3)(*[float(v) for v in np.asarray(bg).reshape(-1)[:3]])
    s.scale_modifier = float(scale_modifier)
    vm = np.ascontiguousarray(np.asarray(viewmatrix, dtype=np.float32).reshape(-1))
    pm = np.ascontiguousarray(np.asarray(projmatrix, dtype=np.float32).reshape(-1))
    s.viewmatrix = (ctypes.c_float * 16)(*vm.tolist())
    s.projmatrix = (ctypes.c_float * 16)(*pm.tolist())
    s.campos = (ctypes.c_float * 3)(*[float(v) for v in np.asarray(campos).reshape(-1)[:3]])
    s.sh_degree = int(sh_degree)
    s.prefiltered = int(bool(prefiltered))
    s.variant = int(variant)
    s.use_filter = int(bool(use_filter))
    s.ogfs_epsilon_scale = float(ogfs_epsilon_scale)
    s.ogfs_epsilon_max = float(ogfs_epsilon_max)
    s.ogfs_energy_floor = float(ogfs_energy_floor)
    s.ogfs_radius_ratio_min = float(ogfs_radius_ratio_min)
    s.grad_mode = int(grad_mode)
    return s


def _dt(mode):
    return np.float32 if mode == "f32" else np.float64


def _c(a, dt):
    if a is None:
        return None
    return np.ascontiguousarray(np.asarray(a, dtype=dt))


def preprocess(s: OrcSettings, means3D, opacities, scales=None, rotations=None, shs=None,
               colors=None, cov3D=None, mode="f32", nthreads=1):
    dt = _dt(mode)
    means3D = _c(means3D, dt).reshape(-1, 3)
    P = means3D.shape[0]
    opac = _c(opacities, dt).reshape(-1)
    sc = _c(scales, dt)
    ro = _c(rotations, dt)
    sh = _c(shs, dt)
    M = 0 if sh is None else sh.reshape(P, -1, 3).shape[1]
    co = _c(colors, dt)
    cv = _c(cov3D, dt)
    out = dict(
        radii=np.zeros(P, np.int32), means2D=np.zeros((P, 2), dt), depths=np.zeros(P, dt),
        cov3D=np.zeros((P, 6), dt), rgb=np.zeros((P, 3), dt), conic_opacity=np.zeros((P, 4), dt),
        tiles_touched=np.zeros(P, np.uint32), clamped=np.zeros((P, 3), np.uint8),
        rect=np.zeros((P, 4), np.int32))
    fn = lib().orc_preprocess_f32 if mode == "f32" else lib().orc_preprocess_f64
    fn(ctypes.byref(s), ctypes.c_int32(P), ctypes.c_int32(M), _p(means3D), _p(sc), _p(ro),
       _p(opac), _p(sh), _p(co), _p(cv), _p(out["radii"]), _p(out["means2D"]), _p(out["depths"]),
       _p(out["cov3D"]), _p(out["rgb"]), _p(out["conic_opacity"]), _p(out["tiles_touched"]),
       _p(out["clamped"]), _p(out["rect"]), ctypes.c_int32(nthreads))
    return out


def ogfs_log_lambda(eps_scale, eps_max) -> float:
    f = lib().orc_ogfs_log_lambda
    f.restype = ctypes.c_float
    return float(f(ctypes.c_float(eps_scale), ctypes.c_float(eps_max)))


def get_rect(px, py, radius, gx, gy):
    out = np.zeros(4, np.int32)
    lib().orc_get_rect(ctypes.c_float(px), ctypes.c_float(py), ctypes.c_int32(radius),
                       ctypes.c_int32(gx), ctypes.c_int32(gy), _p(out))
    return out


def higher_msb(n: int) -> int:
    f = lib().orc_higher_msb
    f.restype = ctypes.c_uint32
    return int(f(ctypes.c_uint32(n)))


def bin_gaussians(radii, rect, depths, W, H):
    """duplicate + stable sort + ranges; returns (keys u64[K], vals u32[K], ranges u32[T,2])."""
    radii = _c(radii, np.int32)
    rect = _c(rect, np.int32)
    depths = _c(depths, np.float32)
    P = radii.shape[0]
    gx, gy = (W + 15) // 16, (H + 15) // 16
    f = lib().orc_num_rendered
    f.restype = ctypes.c_int64
    K = int(f(ctypes.c_int32(P), _p(radii), _p(rect)))
    keys = np.zeros(max(K, 1), np.uint64)
    vals = np.zeros(max(K, 1), np.uint32)
    ranges = np.zeros((gx * gy, 2), np.uint32)
    lib().orc_bin(ctypes.c_int32(P), ctypes.c_int32(gx), ctypes.c_int32(gy), _p(radii), _p(rect),
                  _p(depths), _p(keys), _p(vals), _p(ranges))
    return keys[:K], vals[:K], ranges


def blend_fwd(s: OrcSettings, P, ranges, point_list, means2D, colors, conic_opacity, mode="f32",
              stats=False, decision_hash=False, nthreads=1):
    dt = _dt(mode)
    H, W = s.image_height, s.image_width
    ranges = _c(ranges, np.uint32)
    point_list = _c(point_list, np.uint32)
    if point_list.size == 0:
        point_list = np.zeros(1, np.uint32)
    K = int(ranges.max()) if ranges.size else 0
    out = dict(color=np.zeros((3, H, W), dt), final_T=np.zeros((H, W), dt),
               n_contrib=np.zeros((H, W), np.uint32), point_id=np.zeros((H, W), np.int32),
               weight_pixel=np.zeros((H, W), dt), point_weight=np.zeros(P, dt))
    kv_sum = np.zeros(max(K, 1), dt) if stats else None
    kv_cnt = np.zeros(max(K, 1), np.uint32) if stats else None
    dh = np.zeros((H, W), np.uint64) if decision_hash else None
    fn = lib().orc_blend_fwd_f32 if mode == "f32" else lib().orc_blend_fwd_f64
    fn(ctypes.byref(s), ctypes.c_int32(P), _p(ranges), _p(point_list), _p(_c(means2D, dt)),
       _p(_c(colors, dt)), _p(_c(conic_opacity, dt)), _p(out["color"]), _p(out["final_T"]),
       _p(out["n_contrib"]), _p(out["point_id"]), _p(out["weight_pixel"]),
       _p(out["point_weight"]), _p(kv_sum), _p(kv_cnt), _p(dh), ctypes.c_int32(nthreads))
    if stats:
        out["kv_alpha_sum"] = kv_sum[:K]
        out["kv_alpha_count"] = kv_cnt[:K]
    if decision_hash:
        out["decision_hash"] = dh
    return out


def blend_bwd(s: OrcSettings, P, ranges, point_list, means2D, colors, conic_opacity, final_T,
              n_contrib, dL_dpix, mode="f32", nthreads=1):
    dt = _dt(mode)
    ranges = _c(ranges, np.uint32)
    point_list = _c(point_list, np.uint32)
    if point_list.size == 0:
        point_list = np.zeros(1, np.uint32)
    g = dict(dL_dmean2D=np.zeros((P, 3), dt), dL_dconic=np.zeros((P, 3), dt),
             dL_dopacity=np.zeros(P, dt), dL_dcolors=np.zeros((P, 3), dt))
    fn = lib().orc_blend_bwd_f32 if mode == "f32" else lib().orc_blend_bwd_f64
    fn(ctypes.byref(s), ctypes.c_int32(P), _p(ranges), _p(point_list), _p(_c(means2D, dt)),
       _p(_c(colors, dt)), _p(_c(conic_opacity, dt)), _p(_c(final_T, dt)),
       _p(_c(n_contrib, np.uint32)), _p(_c(dL_dpix, dt)), _p(g["dL_dmean2D"]),
       _p(g["dL_dconic"]), _p(g["dL_dopacity"]), _p(g["dL_dcolors"]), ctypes.c_int32(nthreads))
    return g


def preprocess_bwd(s: OrcSettings, means3D, radii, cov3D, dL_dmean2D, dL_dconic, dL_dopacity,
                   dL_dcolors, scales=None, rotations=None, opacities=None, shs=None, mode="f32",
                   nthreads=1):
    dt = _dt(mode)
    means3D = _c(means3D, dt).reshape(-1, 3)
    P = means3D.shape[0]
    sh = _c(shs, dt)
    M = 0 if sh is None else sh.reshape(P, -1, 3).shape[1]
    dop = _c(dL_dopacity, dt).copy()
    g = dict(dL_dmeans3D=np.zeros((P, 3), dt), dL_dcov3D=np.zeros((P, 6), dt),
             dL_dsh=np.zeros((P, max(M, 0), 3), dt), dL_dscales=np.zeros((P, 3), dt),
             dL_drot=np.zeros((P, 4), dt))
    fn = lib().orc_preprocess_bwd_f32 if mode == "f32" else lib().orc_preprocess_bwd_f64
    fn(ctypes.byref(s), ctypes.c_int32(P), ctypes.c_int32(M), _p(means3D), _p(_c(scales, dt)),
       _p(_c(rotations, dt)), _p(_c(opacities, dt)), _p(sh), _p(_c(cov3D, dt)),
       _p(_c(radii, np.int32)), _p(_c(dL_dmean2D, dt)), _p(_c(dL_dconic, dt)), _p(dop),
       _p(_c(dL_dcolors, dt)), _p(g["dL_dmeans3D"]), _p(g["dL_dcov3D"]),
       _p(g["dL_dsh"] if M > 0 else None), _p(g["dL_dscales"]), _p(g["dL_drot"]),
       ctypes.c_int32(nthreads))
    g["dL_dopacity"] = dop
    return g


def lod_cut(means3d, log_scales, rotations, projmatrix, viewmatrix, tree_node_index,
            ancestor_path, focal_x, focal_y, tanx, tany, padding, radius_max,
            radius_filter_threshold, nthreads=1):
    means3d = _c(means3d, np.float32)
    N = means3d.reshape(-1, 3).shape[0]
    ap = _c(ancestor_path, np.int32).reshape(N, -1)
    L = ap.shape[1]
    mask = np.zeros(N, np.uint8)
    r = np.zeros(N, np.float32)
    d = np.zeros(N, np.float32)
    lib().orc_lod_cut(ctypes.c_int32(N), _p(means3d), _p(_c(log_scales, np.float32)),
                      _p(_c(rotations, np.float32)),
                      _p(_c(np.asarray(projmatrix).reshape(-1), np.float32)),
                      _p(_c(np.asarray(viewmatrix).reshape(-1), np.float32)),
                      _p(_c(tree_node_index, np.int32)), _p(ap), ctypes.c_int32(L),
                      ctypes.c_float(focal_x), ctypes.c_float(focal_y), ctypes.c_float(tanx),
                      ctypes.c_float(tany), ctypes.c_float(padding), ctypes.c_float(radius_max),
                      ctypes.c_float(radius_filter_threshold), _p(mask), _p(r), _p(d),
                      ctypes.c_int32(nthreads))
    return mask.astype(bool), r, d


def radius_v2(means3d, scales, rotations, projmatrix, viewmatrix, fx, fy, tanx, tany):
    means3d = _c(means3d, np.float32)
    P = means3d.reshape(-1, 3).shape[0]
    out = np.zeros(P, np.float32)
    lib().orc_radius_v2(ctypes.c_int32(P), _p(means3d), _p(_c(scales, np.float32)),
                        _p(_c(rotations, np.float32)),
                        _p(_c(np.asarray(projmatrix).reshape(-1), np.float32)),
                        _p(_c(np.asarray(viewmatrix).reshape(-1), np.float32)),
                        ctypes.c_float(fx), ctypes.c_float(fy), ctypes.c_float(tanx),
                        ctypes.c_float(tany), _p(out))
    return out


def log_radius(ids, xyz, scaling, rotation, projmatrix, viewmatrix, fx, fy, tanx, tany):
    """LoG's Gaussian.compute_radius (LoG/model/level_of_gaussian.py:66-88) of
    rows `ids`: raw scaling through a correctly rounded exp, raw rotation
    normalised (eps 1e-12), then compute_radius_cuda
    (LoG/cuda/compute_radius_kernel.cu:107-156) in FP32."""
    ids = _c(np.asarray(ids).reshape(-1), np.int64)
    out = np.zeros(ids.size, np.float32)
    if ids.size:
        lib().orc_log_radius(ctypes.c_int64(ids.size), _p(ids), _p(_c(xyz, np.float32)),
                             _p(_c(scaling, np.float32)), _p(_c(rotation, np.float32)),
                             _p(_c(np.asarray(projmatrix).reshape(-1), np.float32)),
                             _p(_c(np.asarray(viewmatrix).reshape(-1), np.float32)),
                             ctypes.c_float(fx), ctypes.c_float(fy), ctypes.c_float(tanx),
                             ctypes.c_float(tany), _p(out))
    return out


def tree_traverse(node_index, tree, root_index, radius_fn, min_resolution_pixel=3.0,
                  max_depth=1000):
    """TensorTree.traverse (LoG/model/tensor_tree.py:133-188) restated with
    numpy: `radius_fn(ids)` gives the screen radius of rows `ids`.

    Level 0 keeps the roots whose radius is below min_resolution_pixel or that
    have no children; the others form the frontier.  Each further level (1 ..
    min(30, max_depth)) expands the frontier into its children in (frontier
    position, child slot) order, empty slots dropped, keeps the small ones
    and the leaves and continues with the rest; at the limit the frontier
    itself is appended.  Returns int64 ids: kept roots, then each level's kept
    children, then the frontier at the limit."""
    node_index = np.asarray(node_index, np.int64)
    tree = np.asarray(tree, np.int64)
    assert tree.ndim == 2, "tree is [n_nodes, max_child]"
    roots = np.asarray(root_index, np.int64).reshape(-1)
    r = radius_fn(roots)
    keep = (r < min_resolution_pixel) | (node_index[roots] == -1)
    parts = [roots[keep]]
    index = roots[~keep]
    level = 1
    while True:
        if level > 30 or level > max_depth:
            parts.append(index)
            break
        child = tree[node_index[index]].reshape(-1)
        child = child[child != -1]
        r = radius_fn(child)
        keep = (r < min_resolution_pixel) | (node_index[child] == -1)
        parts.append(child[keep])
        if (~keep).sum() == 0:
            break
        index = child[~keep]
        level += 1
    return np.concatenate(parts).astype(np.int64)


def radius_aa(means3d, scales, rotations, viewmatrix, projmatrix, tanx, tany, H, W):
    means3d = _c(means3d, np.float32)
    P = means3d.reshape(-1, 3).shape[0]
    out = np.zeros(P, np.float32)
    lib().orc_radius_aa(ctypes.c_int32(P), _p(means3d), _p(_c(scales, np.float32)),
                        _p(_c(rotations, np.float32)),
                        _p(_c(np.asarray(viewmatrix).reshape(-1), np.float32)),
                        _p(_c(np.asarray(projmatrix).reshape(-1), np.float32)),
                        ctypes.c_float(tanx), ctypes.c_float(tany), ctypes.c_int32(H),
                        ctypes.c_int32(W), _p(out))
    return out


def mark_visible(means3d, viewmatrix, projmatrix):
    means3d = _c(means3d, np.float32)
    P = means3d.reshape(-1, 3).shape[0]
    out = np.zeros(P, np.uint8)
    lib().orc_mark_visible(ctypes.c_int32(P), _p(means3d),
                           _p(_c(np.asarray(viewmatrix).reshape(-1), np.float32)),
                           _p(_c(np.asarray(projmatrix).reshape(-1), np.float32)), _p(out))
    return out.astype(bool)


def visible_flag(xyz, proj, padding=0.05):
    xyz = _c(xyz, np.float32)
    N = xyz.reshape(-1, 3).shape[0]
    valid = np.zeros(N, np.uint8)
    depth = np.zeros(N, np.float32)
    pp = np.zeros((N, 3), np.float32)
    lib().orc_visible_flag(ctypes.c_int32(N), _p(xyz),
                           _p(_c(np.asarray(proj).reshape(-1), np.float32)),
                           ctypes.c_float(padding), _p(valid), _p(depth), _p(pp))
    return valid.astype(bool), depth, pp


# ---------------------------------------------------------------------------
# Convenience: full forward / backward through the oracle.
# ---------------------------------------------------------------------------
def render(s: OrcSettings, means3D, opacities, scales=None, rotations=None, shs=None,
           colors=None, cov3D=None, mode="f32", nthreads=1, stats=False, decision_hash=False):
    """Forward pass.  The tile structure (radii, rectangles, depth keys, sort
    order, ranges) always comes from the f32 emulation; in mode 'f64' the
    continuous quantities (means2D, conic, opacity, colour) and the blend are
    evaluated in double on that structure."""
    pre32 = preprocess(s, means3D, opacities, scales, rotations, shs, colors, cov3D, "f32",
                       nthreads)
    keys, vals, ranges = bin_gaussians(pre32["radii"], pre32["rect"], pre32["depths"],
                                       s.image_width, s.image_height)
    pre = pre32 if mode == "f32" else preprocess(s, means3D, opacities, scales, rotations, shs,
                                                  colors, cov3D, "f64", nthreads)
    P = pre32["radii"].shape[0]
    colors_used = pre["rgb"]
    fwd = blend_fwd(s, P, ranges, vals, pre["means2D"], colors_used, pre["conic_opacity"], mode,
                    stats, decision_hash, nthreads)
    return dict(pre32=pre32, pre=pre, keys=keys, vals=vals, ranges=ranges, **fwd)


def backward(s: OrcSettings, fwd, dL_dpix, means3D, opacities, scales=None, rotations=None,
             shs=None, colors=None, cov3D=None, mode="f32", nthreads=1):
    pre = fwd["pre"]
    P = pre["radii"].shape[0]
    gb = blend_bwd(s, P, fwd["ranges"], fwd["vals"], pre["means2D"], pre["rgb"],
                   pre["conic_opacity"], fwd["final_T"], fwd["n_contrib"], dL_dpix, mode, nthreads)
    gp = preprocess_bwd(s, means3D, fwd["pre32"]["radii"], pre["cov3D"], gb["dL_dmean2D"],
                        gb["dL_dconic"], gb["dL_dopacity"], gb["dL_dcolors"], scales, rotations,
                        opacities, shs, mode, nthreads)
    out = dict(gp)
    out["dL_dmean2D"] = gb["dL_dmean2D"]
    out["dL_dcolors"] = gb["dL_dcolors"]
    out["dL_dconic"] = gb["dL_dconic"]
    return out


# ---------------------------------------------------------------------------
# FilterGS.get_all: gather + Activation.activate_root_return, FP64 numpy.
# ---------------------------------------------------------------------------
_SH_C0 = 0.28209479177387814
_SH_C1 = 0.4886025119029199
_SH_C2 = (1.0925484305920792, -1.0925484305920792, 0.31539156525252005, -1.0925484305920792,
          0.5462742152960396)
_SH_C3 = (-0.5900435899266435, 2.890611442640554, -0.4570457994644658, 0.3731763325901154,
          -0.4570457994644658, 1.445305721320277, -0.5900435899266435)
_SH_C4 = (2.5033429417967046, -1.7701307697799304, 0.9461746957575601, -0.6690465435572892,
          0.10578554691520431, -0.6690465435572892, 0.47308734787878004, -1.7701307697799304,
          0.6258357354491761)


def sh_wobase_basis(dirs, degree):
    """Basis functions 1 .. (degree+1)^2-1 of eval_sh_wobase
    (LoG/model/sh_utils.py:31-68) at unit directions [n,3] -> [n, K]."""
    x, y, z = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    cols = [-_SH_C1 * y, _SH_C1 * z, -_SH_C1 * x]
    if degree > 1:
        xx, yy, zz, xy, yz, xz = x * x, y * y, z * z, x * y, y * z, x * z
        cols += [_SH_C2[0] * xy, _SH_C2[1] * yz, _SH_C2[2] * (2 * zz - xx - yy), _SH_C2[3] * xz,
                 _SH_C2[4] * (xx - yy)]
        if degree > 2:
            cols += [_SH_C3[0] * y * (3 * xx - yy), _SH_C3[1] * xy * z,
                     _SH_C3[2] * y * (4 * zz - xx - yy), _SH_C3[3] * z * (2 * zz - 3 * xx - 3 * yy),
                     _SH_C3[4] * x * (4 * zz - xx - yy), _SH_C3[5] * z * (xx - yy),
                     _SH_C3[6] * x * (xx - 3 * yy)]
            if degree > 3:
                cols += [_SH_C4[0] * xy * (xx - yy), _SH_C4[1] * yz * (3 * xx - yy),
                         _SH_C4[2] * xy * (7 * zz - 1), _SH_C4[3] * yz * (7 * zz - 3),
                         _SH_C4[4] * (zz * (35 * zz - 30) + 3), _SH_C4[5] * xz * (7 * zz - 3),
                         _SH_C4[6] * (xx - yy) * (7 * zz - 1), _SH_C4[7] * xz * (xx - 3 * yy),
                         _SH_C4[8] * (xx * (xx - 3 * yy) - yy * (3 * xx - yy))]
    return np.stack(cols, 1)


def activate_gather(index, xyz, log_scaling, rotation, opacity_logit, features_dc,
                    features_rest, active_sh_degree, campos):
    """FilterGS.get_all (FilterGS/filtergs.py:98-134): row gather, then
    scaling = exp, opacity = sigmoid, rotation = F.normalize (eps 1e-12),
    colors = SH2RGB(dc) + eval_sh_wobase(dir, rest) with dir the normalised
    xyz - campos, no clamp (LoG/model/activation.py:27-43)."""
    idx = np.asarray(index, np.int64)
    f = lambda a: np.asarray(a, np.float64)[idx]
    xyz_g = f(xyz)
    q = f(rotation)
    qn = np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-12)
    colors = f(features_dc).reshape(-1, 3) * _SH_C0 + 0.5
    if active_sh_degree > 0:
        d = xyz_g - np.asarray(campos, np.float64)[None]
        d = d / np.linalg.norm(d, axis=1, keepdims=True)
        B = sh_wobase_basis(d, active_sh_degree)
        rest = f(features_rest).reshape(len(idx), -1, 3)[:, :B.shape[1], :]
        colors = colors + np.einsum("nk,nkc->nc", B, rest)
    return dict(xyz=xyz_g, scaling=np.exp(f(log_scaling)),
                opacity=1.0 / (1.0 + np.exp(-f(opacity_logit).reshape(-1))),
                rotation=q / qn, colors=colors)


def activate_backward(xyz, log_scaling, rotation, opacity_logit, rest_coeffs, active_sh_degree,
                      campos, g_scaling, g_rotation, g_opacity, g_colors):
    """Backward of Activation.activate_root_return (LoG/model/activation.py:
    27-43) for rows in place, in float64: gradients of the raw log-scales,
    quaternions, opacity logits, SH DC and SH rest [n, rest_coeffs, 3] from
    those of the activated scaling, rotation, opacity and colors.  The SH
    direction uses the detached mean (activation.py:30-33), so the mean gets
    no gradient; F.normalize's clamp at 1e-12 stops the norm term below it."""
    f = lambda a: np.asarray(a, np.float64)
    ls, q, z = f(log_scaling), f(rotation), f(opacity_logit).reshape(-1)
    n = len(z)
    d_ls = f(g_scaling) * np.exp(ls)
    g = f(g_rotation)
    nrm = np.linalg.norm(q, axis=1, keepdims=True)
    safe = np.where(nrm >= 1e-12, nrm, 1.0)
    y = q / safe
    d_rot = np.where(nrm >= 1e-12, (g - y * (y * g).sum(1, keepdims=True)) / safe, g / 1e-12)
    o = 1.0 / (1.0 + np.exp(-z))
    d_z = f(g_opacity).reshape(-1) * (1.0 - o) * o
    gc = f(g_colors)
    d_dc = gc * _SH_C0
    d_rest = np.zeros((n, rest_coeffs, 3))
    if active_sh_degree > 0:
        d = f(xyz) - f(campos)[None]
        d = d / np.linalg.norm(d, axis=1, keepdims=True)
        B = sh_wobase_basis(d, active_sh_degree)
        d_rest[:, :B.shape[1], :] = B[:, :, None] * gc[:, None, :]
    return dict(log_scaling=d_ls, rotation=d_rot, opacity_logit=d_z, features_dc=d_dc,
                features_rest=d_rest)


# ---------------------------------------------------------------------------
# Tree formats (apps/ancestor.py), numpy.
# ---------------------------------------------------------------------------
def rebuild_ancestor_path(index_parent, depth, L):
    """apps/ancestor.py:21-81 restated: rows start at -1; level by level
    (1 .. max depth), a node with a parent copies the parent's row and puts
    the parent in column level-1; an orphan (parent -1) gets a row of -1.
    Returns the int32 [N, L] rows; raises IndexError when a level exceeds L
    (the reference's column index would)."""
    index_parent = np.asarray(index_parent, dtype=np.int64)
    depth = np.asarray(depth, dtype=np.int64)
    N = depth.size
    out = np.full((N, L), -1, dtype=np.int32)
    if N == 0 or L == 0:
        return out
    for level in range(1, int(depth.max()) + 1):
        nodes = np.nonzero(depth == level)[0]
        if nodes.size == 0:
            continue
        parents = index_parent[nodes]
        valid = parents >= 0
        if valid.any():
            if level - 1 >= L:
                raise IndexError(f"level {level} exceeds {L} ancestor columns")
            nv, pv = nodes[valid], parents[valid]
            out[nv] = out[pv]
            out[nv, level - 1] = pv.astype(np.int32)
        out[nodes[~valid]] = -1
    return out


# ---------------------------------------------------------------------------
# Training-side neighbours, float32 numpy in the reference's operation order.
# ---------------------------------------------------------------------------
def sparse_adam_step(param, param_sel, grad, exp_avg, exp_avg_sq, index, flag, lr, step,
                     beta1=0.9, beta2=0.999, eps=1e-15, max_exp_avg_sq=None):
    """LoG/model/sparse_optimizer.py:41-78 + 163-195 restated: rows index[j]
    with flag[j] of param / exp_avg / exp_avg_sq (float32 [N, D]) take one
    Adam step from param_sel[j], grad[j].  Returns updated copies
    (param, exp_avg, exp_avg_sq, max_exp_avg_sq)."""
    f32 = np.float32
    p, m_all, v_all = param.copy(), exp_avg.copy(), exp_avg_sq.copy()
    vmax_all = None if max_exp_avg_sq is None else max_exp_avg_sq.copy()
    sel = np.nonzero(np.asarray(flag, bool))[0]
    rows = np.asarray(index)[sel]
    g = grad[sel].astype(f32)
    m = m_all[rows] * f32(beta1) + f32(1.0 - beta1) * g
    v = v_all[rows] * f32(beta2) + f32(1.0 - beta2) * (g * g)
    bc1 = 1.0 - beta1 ** step
    bc2 = 1.0 - beta2 ** step
    vd = v
    if vmax_all is not None:
        vd = np.maximum(vmax_all[rows], v)
        vmax_all[rows] = vd
    denom = np.sqrt(vd) / f32(math.sqrt(bc2)) + f32(eps)
    p[rows] = param_sel[sel] + f32(-(lr / bc1)) * (m / denom)
    m_all[rows] = m
    v_all[rows] = v
    return p, m_all, v_all, vmax_all


def counter_update(counters, visible_index, radii, point_weight, point_id, point_count, grad):
    """LoG/model/counter.py:36-83 restated for one render (distinct rows):
    returns updated copies of the counter dict (create_steps, visible_count,
    weights_max, weights_sum, radii_max, radii_max_max, area_sum, grad_sum)."""
    c = {k: v.copy() for k, v in counters.items()}
    vi = np.asarray(visible_index)
    vis = np.nonzero(np.asarray(radii) > 0)[0]
    rows = vi[vis]
    c["create_steps"][rows] += 1
    c["visible_count"][rows] = (c["visible_count"][rows] + 1).astype(np.int16)
    c["weights_max"][rows] = np.maximum(c["weights_max"][rows], point_weight[vis])
    c["weights_sum"][rows] = c["weights_sum"][rows] + point_weight[vis]
    c["radii_max"][rows] = np.maximum(c["radii_max"][rows], np.asarray(radii)[vis].astype(np.int16))
    prow = vi[np.asarray(point_id)]
    pc = np.asarray(point_count)
    c["area_sum"][prow] = (c["area_sum"][prow] + pc).astype(np.int32)
    if grad is not None:
        gn = np.sqrt(grad[:, 0] * grad[:, 0] + grad[:, 1] * grad[:, 1]).astype(np.float32)
        c["grad_sum"][prow] = c["grad_sum"][prow] + gn[np.asarray(point_id)] * pc.astype(np.float32)
    c["radii_max_max"][prow] = np.maximum(pc.astype(np.int32), c["radii_max_max"][prow])
    return c
