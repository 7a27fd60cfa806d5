"""Benchmark: fwd+bwd of the rasterizer for 3M Gaussians on one MI355X,
batched multi-view vs per-view.

One step = one training-style pass over a batch of 16 cameras (orbit poses at
radius 8-30 around a seeded 3M-Gaussian scene in a 40^3 box, resolutions drawn
from 960x540 / 1920x1080 / 2560x1440 / 3840x2160, SH degree 3, wodilate + OGFS
defaults): forward (preprocess, depth sort, duplicate, tile sort, ranges,
blend) and backward (blend backward, preprocess backward) of every view, with
random dL/dimage.  Inputs are resident in HBM before timing starts.

``value`` is the batched path (one launch per stage for all views, the step
replayed from a HIP graph, blend configuration -- tile layouts, launch order
or persistent launches over a cost-ordered (view, tile) queue -- picked by the
autotuner) in Mpix/s.
Beside it the line reports the same 16 views rendered by 16 sequential plain
per-view calls, and the plain per-view path on one 1920x1080 view of the same
scene with its per-stage times, 16 orbit views all at 1920x1080 batched
against sequential plain calls, and a training step of 3M Gaussians at
1920x1080 (L1 + SSIM through the plugin's autograd, Adam).  The roofline
entry covers the dominant kernel of the batched step (HIP events on the
launch stream around every stage of an eager batched step).  The CPU
baseline is the oracle on a bounded sample of the 1080p view (a 256-row band,
FP32 and FP64, per stage, on the host's threads), whose FP64 image is the
reference of the GPU image's max-abs error on the same band; beside it the
oracle's per-stage times on the CPU configurations (100k at 800x800, 1M at
1080p, 4096 at 128x128 with the finite-difference cost) on one thread and on
all threads.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
HBM_STAGES = ("preprocess", "depth_sort", "tile_offsets", "duplicate", "tile_sort", "ranges",
              "preprocess_bwd")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--gaussians", type=int, default=3_000_000)
    ap.add_argument("--views", type=int, default=16)
    ap.add_argument("--plain-steps", type=int, default=20, help="steps of the plain 1080p leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-configs", action="store_true",
                    help="skip the oracle's per-stage timings of the CPU configurations")
    ap.add_argument("--no-train", action="store_true", help="skip the L1+SSIM training-step leg")
    ap.add_argument("--no-1080p-batch", action="store_true", help="skip the 16x1080p batched leg")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-lod", action="store_true", help="skip the urban LoD leg")
    ap.add_argument("--cpu-band", type=int, default=256, help="rows of the CPU baseline sample")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after timing, run one device-to-device copy of this many MiB (a "
                         "reference kernel of known traffic for hardware byte counters)")
    return ap.parse_args()


def stage_bytes(P, Pv, K, T, npix, M, passes_tile, V=1):
    """Algorithmic HBM bytes per launch of each stage.  P Gaussians, V views,
    Pv visible (Gaussian, view) pairs, K kv pairs, T tiles and npix pixels over
    all views.  The batched preprocess reads the Gaussians once for all views;
    the plain depth sort makes 4 passes over all P keys (culled ones sort
    last), the batched one 5 passes over the Pv visible keys only (the
    preprocess appends just those).  The batched duplication is fused with
    the tile sort's first pass: its counting kernel and its scatter each read
    the depth-ordered records (20 B per visible pair) and the scatter writes
    the pairs once (8 B), so its tile sort stage holds the other passes."""
    params = P * (44 + 12 * M)
    grads_out = P * (92 + 12 * M)
    sorted_keys = P if V == 1 else Pv
    out = {
        "preprocess": params + V * P * 8 + Pv * 40,
        "depth_sort": (4 if V == 1 else 5) * 24 * sorted_keys,
        "tile_offsets": 8 * sorted_keys,
        "duplicate": (20 if V == 1 else 40) * Pv + 8 * K,
        "tile_sort": (passes_tile if V == 1 else passes_tile - 1) * 24 * K,
        "ranges": 8 * K + 8 * T,
        "blend_fwd": 40 * K + 28 * npix + 4 * V * P,
        "blend_bwd": 40 * K + 20 * npix + 36 * Pv,
        "preprocess_bwd": (Pv * (44 + 12 * M + 36) if V == 1 else params + 36 * Pv) + grads_out,
    }
    # the blend kernels alone (their stages also hold the pair-mark clear and
    # the gradient gather): the same algorithmic bytes
    out["blend_fwd_kernel"] = out["blend_fwd"]
    out["blend_bwd_kernel"] = out["blend_bwd"]
    return out


def filtergs_views(tree, cams, warmup):
    """FilterGS inference on the LoD tree (FilterGS/filtergs.py:20-223): per
    view prepare() (the cut, radius_max 5.0 as render.yml) and render(), with
    render() once folded into the rasterizer (rows activated in the
    preprocess) and once on the gathered route of the reference (gather +
    activation, finite filter, rasterizer), forward only, statistics on."""
    import torch

    from filtergs_amd.filtergs import FilterGS
    from filtergs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer
    dev = "cuda"
    fg = FilterGS(sh_degree=1)
    fg.register_buffer("xyz", tree["means3D"])
    fg.register_buffer("scaling", tree["log_scales"])
    fg.register_buffer("rotation", tree["rotations"])
    op = tree["opacities"].clamp(1e-4, 1 - 1e-4)
    fg.register_buffer("opacity", torch.log(op / (1 - op)))
    fg.register_buffer("colors", tree["shs"][:, 0, :].contiguous())
    fg.register_buffer("shs", tree["shs"][:, 1:, :].contiguous())
    fg.active_sh_degree = 1

    class Gathered:
        def __init__(self, r):
            self.r, self.raster_settings = r, r.raster_settings

        def __call__(self, **kw):
            return self.r(**kw)

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    out = {"prepare_ms": [], "render_rows_ms": [], "render_gathered_ms": [], "rendered": []}
    t = lambda a: torch.as_tensor(a, dtype=torch.float32, device=dev)
    for cam in cams:
        rs = GaussianRasterizationSettings(
            image_height=cam.height, image_width=cam.width, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy,
            bg=t([0.0, 0.0, 0.0]), scale_modifier=1.0, viewmatrix=t(cam.world_view_transform),
            projmatrix=t(cam.full_proj_transform), sh_degree=1, campos=t(cam.center), prefiltered=False,
            debug=False)
        r = GaussianRasterizer(rs)
        camera = {"camera_center": t(cam.center), "full_proj_transform": t(cam.full_proj_transform)}

        def run(fn):
            for _ in range(max(1, warmup)):
                fn()
            torch.cuda.synchronize()
            ev[0].record()
            fn()
            ev[1].record()
            ev[1].synchronize()
            return round(ev[0].elapsed_time(ev[1]), 3)

        prep = lambda: fg.prepare(r, camera, tree_node_index_gpu=tree["tree_node_index"],
                                  ancestor_path_gpu=tree["ancestor_path"], radius_max=5.0,
                                  padding=0.1, radius_filter_threshold=700.0)
        out["prepare_ms"].append(run(prep))
        out["rendered"].append(int(fg.visibility_flag["index"].numel()))
        out["render_rows_ms"].append(run(lambda: fg.render(camera, r)))
        g = Gathered(r)
        out["render_gathered_ms"].append(run(lambda: fg.render(camera, g)))
    n = len(cams)
    out["mean"] = {k: round(sum(out[k]) / n, 3) for k in ("prepare_ms", "render_rows_ms", "render_gathered_ms")}
    return out


def traverse_views(tree, cams, min_px=3.0):
    """LoG's TensorTree.traverse (LoG/model/tensor_tree.py:133-188) of the
    LoD tree in TensorTree layout from each aerial view, on the device, and
    the reference's level loop in torch for comparison (per level: children
    gathered, radius of exp(scaling) / normalize(rotation) by the
    compute_radius kernel, keep/next masks, a host wait on the next-level
    count).  Wall times: both wait for their result."""
    import time

    import torch

    from filtergs_amd import lod, workloads
    from filtergs_amd.tensor_tree import tree_traverse
    dev = "cuda"
    node_index, rows = workloads.children_table(tree["index_parent"], 4)
    roots = torch.nonzero(tree["index_parent"] < 0).reshape(-1)
    xyz, lsc, rot = tree["means3D"], tree["log_scales"], tree["rotations"]
    hint = int(tree["level"].max().item())
    ni64 = node_index.long()

    def torch_loop(cam, fx, fy, pm, vm):
        def radius(ids):
            return lod.compute_radius_module.compute_radius(
                xyz[ids], torch.exp(lsc[ids]), torch.nn.functional.normalize(rot[ids]), pm, vm, fx, fy,
                cam.tanfovx, cam.tanfovy)
        r = radius(roots)
        keep = (r < min_px) | (ni64[roots] == -1)
        parts = [roots[keep]]
        index = roots[~keep]
        level = 1
        while True:
            if level > 30:
                parts.append(index)
                break
            child = rows[ni64[index]].reshape(-1).long()
            child = child[child != -1]
            r = radius(child)
            keep = (r < min_px) | (ni64[child] == -1)
            parts.append(child[keep])
            go = ~keep
            if go.sum() == 0:
                break
            index = child[go]
            level += 1
        return torch.cat(parts)

    out = {"device_ms": [], "torch_loop_ms": [], "ids": [], "equal": [], "differing_ids": []}
    for cam in cams:
        fx = cam.width / (2 * cam.tanfovx)
        fy = cam.height / (2 * cam.tanfovy)
        pm = torch.as_tensor(cam.full_proj_transform, dtype=torch.float32, device=dev)
        vm = torch.as_tensor(cam.world_view_transform, dtype=torch.float32, device=dev)
        dev_fn = lambda: tree_traverse(xyz, lsc, rot, node_index, rows, roots, vm, pm, fx, fy, cam.tanfovx,
                                       cam.tanfovy, min_px, 1000, level_hint=hint)
        ref_fn = lambda: torch_loop(cam, fx, fy, pm, vm)
        res = {}
        for name, fn in (("device_ms", dev_fn), ("torch_loop_ms", ref_fn)):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res[name] = fn()
            torch.cuda.synchronize()
            out[name].append(round((time.perf_counter() - t0) * 1e3, 3))
        a, b = res["device_ms"], res["torch_loop_ms"]
        out["ids"].append(int(a.numel()))
        out["equal"].append(bool(a.numel() == b.numel() and torch.equal(a, b)))
        sa, sb = set(a.tolist()), set(b.tolist())
        out["differing_ids"].append(len(sa ^ sb))
    n = len(cams)
    out["mean"] = {k: round(sum(out[k]) / n, 3) for k in ("device_ms", "torch_loop_ms")}
    return out


def lod_leg(steps, warmup):
    """Urban-scale LoD: a 22.4M-node quadtree (12 levels, 2x2 roots over 2 km x
    2 km, leaves on a seeded height field, SH degree 1), the LoD cut of each of
    8 oblique aerial 1920x1080 views (radius_max 1.5 px, padding 0.1,
    threshold 700), the attribute gather of the union of the cuts, then the
    batched fwd+bwd of the 8 views over the gathered Gaussians."""
    import torch

    from filtergs_amd import lod, workloads
    from filtergs_amd.batch import BatchRasterizer
    dev = "cuda"
    tree = workloads.lod_tree(levels=12, roots_side=2, extent=2000.0, sh_degree=1, seed=0, device=dev)
    cams = workloads.aerial_cameras(8)
    N = int(tree["means3D"].shape[0])
    L = int(tree["ancestor_path"].shape[1])
    n_leaf = int((tree["tree_node_index"] < 0).sum().item())
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]

    def cut(cam, rmax):
        f = cam.width / (2 * cam.tanfovx)
        return lod.compute_visibility_radius(
            tree["means3D"], tree["log_scales"], tree["rotations"],
            torch.as_tensor(cam.full_proj_transform, device=dev),
            torch.as_tensor(cam.world_view_transform, device=dev), tree["tree_node_index"],
            tree["ancestor_path"], f, cam.height / (2 * cam.tanfovy), cam.tanfovx, cam.tanfovy,
            rmax, padding=0.1, radius_filter_threshold=700.0)

    def cut_parent(cam, rmax):
        f = cam.width / (2 * cam.tanfovx)
        return lod.compute_visibility_radius_parent(
            tree["means3D"], tree["log_scales"], tree["rotations"],
            torch.as_tensor(cam.full_proj_transform, device=dev),
            torch.as_tensor(cam.world_view_transform, device=dev), tree["tree_node_index"],
            tree["index_parent"], f, cam.height / (2 * cam.tanfovy), cam.tanfovx, cam.tanfovy,
            rmax, padding=0.1, radius_filter_threshold=700.0)

    res = {}
    for rmax in (5.0, 1.5):
        cut(cams[0], rmax)
        torch.cuda.synchronize()
        ev[0].record()
        masks = [cut(c, rmax)[0] for c in cams]
        ev[1].record()
        ev[1].synchronize()
        ms = ev[0].elapsed_time(ev[1]) / len(cams)
        cut_parent(cams[0], rmax)
        torch.cuda.synchronize()
        ev[0].record()
        masks_p = [cut_parent(c, rmax)[0] for c in cams]
        ev[1].record()
        ev[1].synchronize()
        ms_p = ev[0].elapsed_time(ev[1]) / len(cams)
        if not all(torch.equal(a, b) for a, b in zip(masks, masks_p)):
            raise RuntimeError("parent-link cut differs from the ancestor-row cut")
        union = torch.stack(masks).any(0)
        nbytes = N * (44 + 4 + 9) + n_leaf * 4 * L
        nbytes_p = N * (44 + 4 + 9) + n_leaf * 4
        res[rmax] = {"ms_per_view": round(ms, 3), "GB/s": round(nbytes / ms / 1e6, 1),
                     "parent_links_ms_per_view": round(ms_p, 3),
                     "parent_links_GB/s": round(nbytes_p / ms_p / 1e6, 1),
                     "kept_per_view": [int(m.sum().item()) for m in masks],
                     "union": int(union.sum().item())}
    # the ancestor rows rebuilt on the GPU from the parent links (apps/ancestor.py)
    sd = {"tree.index_parent": tree["index_parent"], "tree.depth": tree["level"],
          "tree.ancestor_path": tree["ancestor_path"]}
    lod.rebuild_ancestor_path(dict(sd))
    torch.cuda.synchronize()
    ev[0].record()
    rows, _ = lod.rebuild_ancestor_path(dict(sd))
    ev[1].record()
    ev[1].synchronize()
    ms_rebuild = ev[0].elapsed_time(ev[1])
    if not torch.equal(rows, tree["ancestor_path"]):
        raise RuntimeError("rebuilt ancestor rows differ")
    del rows
    fg_res = filtergs_views(tree, cams, warmup)
    trav_res = traverse_views(tree, cams)
    union = torch.stack([cut(c, 1.5)[0] for c in cams]).any(0)
    idx = torch.nonzero(union).reshape(-1)
    torch.cuda.synchronize()
    ev[0].record()
    ga = lod.gather_attributes(idx, tree["means3D"], tree["log_scales"], tree["rotations"],
                               tree["opacities"], tree["means3D"], tree["shs"])
    ev[1].record()
    ev[1].synchronize()
    ms_gather = ev[0].elapsed_time(ev[1])
    P = int(idx.numel())
    means = ga["xyz"].contiguous()
    scales = ga["scaling"].exp().contiguous()
    rots = torch.nn.functional.normalize(ga["rotation"], dim=1).contiguous()
    opac = ga["opacity"].reshape(-1).contiguous()
    shs = ga["shs"].contiguous()
    del tree
    br = BatchRasterizer(cams, P, bg=(0.0, 0.0, 0.0), sh_degree=1)
    br.bind(means, opac, scales, rots, shs=shs)
    K = br.forward()
    npix = int(br.sizes.pixels)
    dL = torch.randn((3 * npix,), device=dev, generator=torch.Generator(device=dev).manual_seed(4))
    grads = br.backward(dL)
    graph, _ = br.capture(dL, out=grads)
    for _ in range(warmup):
        graph.replay()
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(steps):
        graph.replay()
    ev[1].record()
    ev[1].synchronize()
    ms = ev[0].elapsed_time(ev[1]) / steps
    if int(br.counts[1].item()) > br.K_cap:
        raise RuntimeError("kv pairs exceeded the reserved capacity during replay")
    return {"nodes": N, "leaves": n_leaf, "levels": L,
            "cut_radius_max_1.5": res[1.5], "cut_radius_max_5.0": res[5.0],
            "filtergs_render": fg_res, "tensor_tree_traverse": trav_res,
            "rebuild_ancestor_ms": round(ms_rebuild, 3),
            "gather_ms": round(ms_gather, 3), "gaussians_rendered": P, "kv_pairs": K,
            "batched_fwd_bwd_ms": round(ms, 3), "Mpix/s": round(npix / ms / 1000.0, 3),
            "views": len(cams), "pixels": npix}


def train_leg(g_t, cam, steps, warmup):
    """Training step of 3M Gaussians at 1920x1080 through the plugin's
    autograd (LoG/render/renderer.py:260-273): render, 0.8 L1 + 0.2 (1 -
    SSIM) against a seeded target, backward to the five parameter tensors and
    means2D, and a dense Adam step over them; timed with and without the
    optimizer step."""
    import torch

    from filtergs_amd import workloads
    from filtergs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer
    dev = "cuda"
    t = lambda a: torch.as_tensor(a, dtype=torch.float32, device=dev)
    H, W = cam.height, cam.width
    rs = GaussianRasterizationSettings(
        image_height=H, image_width=W, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy,
        bg=t([0.0, 0.0, 0.0]), scale_modifier=1.0, viewmatrix=t(cam.world_view_transform),
        projmatrix=t(cam.full_proj_transform), sh_degree=3, campos=t(cam.center), prefiltered=False,
        debug=False)
    r = GaussianRasterizer(rs)
    means, scales, rots, opac, shs = (x.detach().clone().requires_grad_(True) for x in g_t)
    means2D = torch.zeros_like(means, requires_grad=True)
    params = [means, scales, rots, opac, shs]
    opt = torch.optim.Adam(params, lr=1e-6, foreach=True)
    target = torch.rand((3, H, W), generator=torch.Generator(device=dev).manual_seed(5), device=dev)
    win = workloads.ssim_window().to(dev)

    def step(with_opt):
        for p in params + [means2D]:
            p.grad = None
        color = r(means3D=means, means2D=means2D, opacities=opac, shs=shs, scales=scales,
                  rotations=rots)[0]
        loss = workloads.l1_ssim_loss(color, target, win)
        loss.backward()
        if with_opt:
            opt.step()
        return loss

    res = {}
    for name, wo in (("render_loss_backward_ms", False), ("step_with_adam_ms", True)):
        for _ in range(max(1, warmup)):
            step(wo)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            loss = step(wo)
        torch.cuda.synchronize()
        res[name] = round((time.perf_counter() - t0) / steps * 1e3, 3)
    res["loss"] = round(float(loss.detach()), 6)
    res["Mpix/s"] = round(H * W / res["render_loss_backward_ms"] / 1e3, 3)
    res["workload"] = f"{means.shape[0]} Gaussians, {W}x{H}, SH3, 0.8 L1 + 0.2 (1 - SSIM 11x11)"
    return res


def batch_1080p_leg(g_t, views, steps, warmup, plain_step):
    """`views` orbit cameras at 1920x1080 over the same scene: the batched
    fwd+bwd (graph replay) against the views rendered by sequential plain
    per-view calls."""
    import torch

    from filtergs_amd import workloads
    from filtergs_amd.batch import BatchRasterizer
    dev = "cuda"
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float32, device=dev)
    means, scales, rots, opac, shs = g_t
    cams = workloads.orbit_cameras(views, seed=5, radius_range=(8.0, 30.0), resolutions=((1920, 1080),))
    rng = np.random.default_rng(6)
    dLs = [t(rng.normal(size=(3, c.height, c.width)).astype(np.float32)) for c in cams]
    cts = [(t(c.world_view_transform), t(c.full_proj_transform), t(c.center)) for c in cams]
    bg0 = t([0.0, 0.0, 0.0])

    def seq():
        for c, ct, d in zip(cams, cts, dLs):
            plain_step(c, ct, d, bg0)

    def timed(fn, n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3

    seq()
    ms_seq = timed(seq, max(2, steps // 2))
    br = BatchRasterizer(cams, means.shape[0], bg=(0.0, 0.0, 0.0))
    br.bind(means, opac.reshape(-1), scales, rots, shs=shs)
    K = br.forward()
    dL = torch.cat([d.reshape(-1) for d in dLs])
    grads = br.backward(dL)
    graph, _ = br.capture(dL, out=grads)
    for _ in range(warmup):
        graph.replay()
    ms = timed(graph.replay, steps)
    if int(br.counts[1].item()) > br.K_cap:
        raise RuntimeError("kv pairs exceeded the reserved capacity during replay")
    npix = views * 1920 * 1080
    out = {"views": views, "kv_pairs": K, "batched_ms": round(ms, 3),
           "batched_ms_per_view": round(ms / views, 3), "batched_Mpix/s": round(npix / ms / 1e3, 3),
           "sequential_ms": round(ms_seq, 3), "sequential_ms_per_view": round(ms_seq / views, 3),
           "batched_speedup": round(ms_seq / ms, 3)}
    del br, graph, grads
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_stages(orc, s, a, mode, nthreads, dL=None):
    """Wall time of each oracle stage for one view, in ms: preprocess (the
    FP32 tile structure, plus the FP64 attributes in mode f64), binning
    (duplicate, stable sort, ranges), blend forward and, when dL is given,
    blend backward and preprocess backward.  Inputs `a` are already in the
    mode's type.  Returns (stage dict, forward result, backward result)."""
    t = {}
    t0 = time.perf_counter()
    p32 = orc.preprocess(s, *a["f32"], mode="f32", nthreads=nthreads)
    pre = p32 if mode == "f32" else orc.preprocess(s, *a["f64"], mode="f64", nthreads=nthreads)
    t["preprocess"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    keys, vals, ranges = orc.bin_gaussians(p32["radii"], p32["rect"], p32["depths"], s.image_width,
                                           s.image_height)
    t["binning"] = time.perf_counter() - t0
    P = p32["radii"].shape[0]
    t0 = time.perf_counter()
    f = orc.blend_fwd(s, P, ranges, vals, pre["means2D"], pre["rgb"], pre["conic_opacity"], mode,
                      decision_hash=True, nthreads=nthreads)
    t["blend_fwd"] = time.perf_counter() - t0
    gp = None
    if dL is not None:
        m = a[mode]
        t0 = time.perf_counter()
        gb = orc.blend_bwd(s, P, ranges, vals, pre["means2D"], pre["rgb"], pre["conic_opacity"],
                           f["final_T"], f["n_contrib"], dL, mode, nthreads=nthreads)
        t["blend_bwd"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        gp = orc.preprocess_bwd(s, m[0], p32["radii"], pre["cov3D"], gb["dL_dmean2D"], gb["dL_dconic"],
                                gb["dL_dopacity"], gb["dL_dcolors"], m[2], m[3], m[1], m[4], mode,
                                nthreads=nthreads)
        t["preprocess_bwd"] = time.perf_counter() - t0
    out = {k: round(v * 1e3, 2) for k, v in t.items()}
    out["fwd"] = round(sum(t[k] for k in ("preprocess", "binning", "blend_fwd")) * 1e3, 2)
    if dL is not None:
        out["bwd"] = round((t["blend_bwd"] + t["preprocess_bwd"]) * 1e3, 2)
    out["kv_pairs"] = int(len(vals))
    return out, f, gp


def _typed(g):
    keys = ("means3D", "opacities", "scales", "rotations", "shs")
    return {m: tuple(np.ascontiguousarray(g[k], dt) for k in keys)
            for m, dt in (("f32", np.float32), ("f64", np.float64))}


def cpu_configs(cores):
    """The oracle per stage on the CPU configurations, FP64 and FP32, on one
    thread and on `cores` threads: 0 -- 100k Gaussians at 800x800, dilate
    variant, forward; 1 -- 1M Gaussians at 1920x1080, forward and backward
    (full frame on all threads, a 64-row band on one thread: per-Mpix rates);
    2a -- 4096 Gaussians at 128x128, forward and backward, and the FP64
    central finite difference of one parameter (two forwards on one thread),
    extrapolated to all 59 parameters of every Gaussian."""
    from filtergs_amd import workloads
    from oracle import pyoracle as orc
    out = {}

    def settings(cam, variant=0):
        return orc.settings(cam.height, cam.width, cam.tanfovx, cam.tanfovy, cam.world_view_transform,
                            cam.full_proj_transform, cam.center, sh_degree=3, variant=variant)

    def entry(st, cam, threads, rows=None):
        tot = st["fwd"] + st.get("bwd", 0.0)
        return {"threads": threads, "rows": rows or cam.height, "stages_ms": st,
                "Mpix/s": round(cam.width * cam.height / tot / 1e3, 4)}

    # 0: forward only, dilate
    g, cam = workloads.config("c0")
    a = _typed(g)
    s = settings(cam, variant=1)
    out["c0_100k_800x800_dilate_fwd"] = {
        m: [entry(oracle_stages(orc, s, a, m, n)[0], cam, n) for n in (1, cores)] for m in ("f64", "f32")}
    # 1: 1M at 1080p
    g, cam = workloads.config("c1")
    a = _typed(g)
    s = settings(cam)
    rng = np.random.default_rng(11)
    dL = rng.normal(size=(3, cam.height, cam.width))
    rows = 64
    y0 = (cam.height - rows) // 2
    bcam = workloads.band_camera(cam, y0, rows)
    bs = settings(bcam)
    res = {}
    for m in ("f64", "f32"):
        d = np.ascontiguousarray(dL, np.float64 if m == "f64" else np.float32)
        full = entry(oracle_stages(orc, s, a, m, cores, d)[0], cam, cores)
        band = entry(oracle_stages(orc, bs, a, m, 1, np.ascontiguousarray(d[:, y0:y0 + rows]))[0], bcam,
                     1, rows)
        band["ms_per_view_extrapolated"] = round(
            band["stages_ms"]["preprocess"] + band["stages_ms"]["preprocess_bwd"] +
            (band["stages_ms"]["binning"] + band["stages_ms"]["blend_fwd"] +
             band["stages_ms"]["blend_bwd"]) * cam.height / rows, 1)
        band["Mpix/s"] = round(cam.width * cam.height / band["ms_per_view_extrapolated"] / 1e3, 4)
        res[m] = [band, full]
    out["c1_1M_1080p_fwd_bwd"] = res
    del g, a, dL
    # 2a: 4096 at 128x128, and the finite-difference cost
    g, cam = workloads.config("c2a")
    a = _typed(g)
    s = settings(cam)
    dL = np.random.default_rng(12).normal(size=(3, cam.height, cam.width))
    oracle_stages(orc, s, a, "f32", cores)      # thread pool start-up, untimed
    res = {m: [entry(oracle_stages(orc, s, a, m, n,
                                   np.ascontiguousarray(dL, np.float64 if m == "f64" else np.float32))[0],
                     cam, n) for n in (1, cores)] for m in ("f64", "f32")}
    m64 = list(a["f64"])
    n_fd = 16
    t0 = time.perf_counter()
    for j in range(n_fd):
        i = j * 37 % m64[0].shape[0]
        for sgn in (1.0, -1.0):
            mm = m64[0].copy()
            mm[i, 0] += sgn * 1e-6
            orc.render(s, mm, *m64[1:], mode="f64", nthreads=1)
    per = (time.perf_counter() - t0) / n_fd
    n_par = g["means3D"].shape[0] * (3 + 1 + 3 + 4 + 3 * g["shs"].shape[1])
    res["fd_f64"] = {"ms_per_parameter": round(per * 1e3, 2), "parameters": n_par,
                     "all_parameters_s_extrapolated": round(per * n_par, 1), "threads": 1}
    out["c2a_4096_128x128"] = res
    return out


def cpu_baseline(g, cam, dL, band, gpu_band):
    """The oracle on a horizontal band of the 3M-Gaussian 1080p view (all
    Gaussians preprocessed, the band's pixels blended forward and backward)
    on all host threads: FP32 emulation (the reported value) and FP64, per
    stage.  The FP64 band image is also the reference for the GPU image of
    the same band (max-abs on pixels where the FP32 and FP64 oracles and the
    GPU take the same alpha/transmittance decisions; the rest counted)."""
    from filtergs_amd import workloads
    from oracle import pyoracle as orc
    y0 = (cam.height - band) // 2
    sub = workloads.band_camera(cam, y0, band)
    s = orc.settings(band, cam.width, sub.tanfovx, sub.tanfovy, sub.world_view_transform,
                     sub.full_proj_transform, sub.center, sh_degree=3)
    cores = min(os.cpu_count() or 1, 16)
    a = _typed(g)
    d32 = np.ascontiguousarray(dL[:, y0:y0 + band, :], np.float32)
    st32, f32, _ = oracle_stages(orc, s, a, "f32", cores, d32)
    st64, f64, _ = oracle_stages(orc, s, a, "f64", cores, d32.astype(np.float64))
    tot = (st32["fwd"] + st32["bwd"]) / 1e3
    same = (f64["decision_hash"] == f32["decision_hash"])
    gd32 = np.abs(gpu_band - f32["color"]).max(0)
    agree = same & (gd32 <= 2e-5)
    d64 = np.abs(gpu_band - f64["color"]).max(0)
    accuracy = {"pixels": int(agree.size), "maxabs_vs_f64": float(d64[agree].max()),
                "maxabs_vs_f64_all_pixels": float(d64.max()),
                "decision_flips": int((~agree).sum()),
                "maxabs_vs_f32_emulation": float(gd32.max()),
                "sample": f"rows {y0}..{y0 + band} of the {cam.width}x{cam.height} view, "
                          f"{g['means3D'].shape[0]} Gaussians"}
    out = {"value": round(cam.width * band / tot / 1e6, 4), "unit": "Mpix/s", "cores": cores,
           "kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
           "sample": f"oracle FP32 emulation fwd+bwd of rows {y0}..{y0 + band} of one "
                     f"{cam.width}x{cam.height} view, all {g['means3D'].shape[0]} Gaussians "
                     f"preprocessed, every stage on {cores} threads, {tot:.1f} s",
           "stages_ms_f32": st32, "stages_ms_f64": st64,
           "f64_Mpix/s": round(cam.width * band / (st64["fwd"] + st64["bwd"]) / 1e3, 4)}
    return out, accuracy


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch

    from filtergs_amd import _lib, ext, workloads
    from filtergs_amd.batch import BatchRasterizer
    P = args.gaussians
    g = workloads.gaussians(P, 20.0, 3, seed=0)
    dev = "cuda"

    def t(a, dt=torch.float32):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)

    means, scales, rots = t(g["means3D"]), t(g["scales"]), t(g["rotations"])
    opac, shs = t(g["opacities"].reshape(-1, 1)), t(g["shs"])
    empty = torch.Tensor([])
    M = 16

    def plain_step(cam, cams_t, dLt, bg):
        vm, pm, cp = cams_t
        out = ext.rasterize_gaussians(bg, means, empty, opac, scales, rots, 1.0, empty, vm, pm,
                                      cam.tanfovx, cam.tanfovy, cam.height, cam.width, shs, 3, cp,
                                      False, True, False, ogfs_enable_stats=False)
        K, color, pid, wpix, pw, radii, geom, binning, img, _ = out
        ext.rasterize_gaussians_backward(bg, means, radii, empty, scales, rots, 1.0, empty, vm, pm,
                                         cam.tanfovx, cam.tanfovy, dLt, shs, 3, cp, geom, K,
                                         binning, img, False)
        return K, radii

    def timed(fn, steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    def stages_of(prof, nbytes):
        out = {}
        for name, (tot, calls) in prof.items():
            if calls:
                avg = tot / calls
                out[name] = {"ms": round(avg, 4), "GB/s": round(nbytes[name] / avg / 1e6, 1)}
        return out

    def hbm_frac(st, nbytes):
        b = sum(nbytes[k] for k in HBM_STAGES if k in st)
        ms = sum(st[k]["ms"] for k in HBM_STAGES if k in st)
        return {"GB/s": round(b / ms / 1e6, 1), "frac": round(b / ms / 1e6 / HBM_PEAK_GBS, 4)}

    bg0 = t([0.0, 0.0, 0.0])

    # ---- plain per-view path, one 1920x1080 view (stage breakdown) --------
    cam1 = workloads.camera_c1(1920, 1080)
    ct1 = (t(cam1.world_view_transform), t(cam1.full_proj_transform), t(cam1.center))
    dL1 = np.random.default_rng(1).normal(size=(3, 1080, 1920)).astype(np.float32)
    dL1t = t(dL1)
    for _ in range(max(1, args.warmup)):
        K1, radii1 = plain_step(cam1, ct1, dL1t, bg0)
    torch.cuda.synchronize()
    Pv1 = int((radii1 > 0).sum().item())
    _lib.profile_enable(True)
    _lib.profile_read(reset=True)
    ms1 = timed(lambda: plain_step(cam1, ct1, dL1t, bg0), args.plain_steps) * 1e3
    prof1 = _lib.profile_read(reset=True)
    _lib.profile_enable(False)
    T1 = 120 * 68
    nb1 = stage_bytes(P, Pv1, K1, T1, 1920 * 1080, M, (T1.bit_length() + 7) // 8)
    st1 = stages_of(prof1, nb1)
    # the GPU image of the band the CPU baseline renders (its FP64 reference)
    yb = (cam1.height - args.cpu_band) // 2
    bcam = workloads.band_camera(cam1, yb, args.cpu_band)
    bct = (t(bcam.world_view_transform), t(bcam.full_proj_transform), t(bcam.center))
    gpu_band = ext.rasterize_gaussians(bg0, means, empty, opac, scales, rots, 1.0, empty, bct[0], bct[1],
                                       bcam.tanfovx, bcam.tanfovy, bcam.height, bcam.width, shs, 3,
                                       bct[2], False, True, False, ogfs_enable_stats=False)[1]
    gpu_band = gpu_band.cpu().numpy()

    # ---- the views: 16 sequential plain calls ------------------------------
    cams = workloads.orbit_cameras(args.views, seed=3, radius_range=(8.0, 30.0))
    npix = sum(c.width * c.height for c in cams)
    rng = np.random.default_rng(2)
    dLs = [t(rng.normal(size=(3, c.height, c.width)).astype(np.float32)) for c in cams]
    cts = [(t(c.world_view_transform), t(c.full_proj_transform), t(c.center)) for c in cams]

    def seq():
        for c, ct, d in zip(cams, cts, dLs):
            plain_step(c, ct, d, bg0)

    seq()
    ms_seq = timed(seq, max(2, args.steps // 2)) * 1e3

    # ---- batched: autotune, eager stage profile, graph replay --------------
    br = BatchRasterizer(cams, P, bg=(0.0, 0.0, 0.0))
    br.bind(means, opac.reshape(-1), scales, rots, shs=shs)
    K = br.forward()
    dL_flat = torch.cat([d.reshape(-1) for d in dLs])
    grads = br.backward(dL_flat)
    tuned = None if args.no_autotune else br.autotune(dL_flat)
    Pv = int((br.radii > 0).sum().item())

    def eager():
        br._forward_launch()
        br._backward_launch(dL_flat, grads)

    _lib.profile_enable(True)
    _lib.profile_read(reset=True)
    ms_eager = timed(eager, 3) * 1e3
    profb = _lib.profile_read(reset=True)
    _lib.profile_enable(False)
    graph, _ = br.capture(dL_flat, out=grads)
    for _ in range(args.warmup):
        graph.replay()
    ms_graph = timed(graph.replay, args.steps) * 1e3
    if int(br.counts[1].item()) > br.K_cap:
        raise RuntimeError("kv pairs exceeded the reserved capacity during replay")
    T = int(br.sizes.tiles)
    # the batch's tile keys hold the view's own tile number: the largest view
    # sets the tile sort's passes
    nbb = stage_bytes(P, Pv, K, T, npix, M, (int(br.sizes.max_view_tiles).bit_length() + 7) // 8,
                      V=args.views)
    stb = stages_of(profb, nbb)
    dominant = max((k for k in stb if not k.endswith("_kernel")), key=lambda k: stb[k]["ms"])
    # the kernel's own launch time where the stage holds more than the kernel
    kern = stb.get(dominant + "_kernel", stb[dominant])
    achieved = nbb[dominant] / kern["ms"] / 1e6
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dominant}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")

    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    res = sorted({f"{c.width}x{c.height}" for c in cams})
    line = {
        "metric": "fwd+bwd ms/view & Mpix/s, 3M Gaussians @1080p; batched vs per-view; % HBM peak",
        "value": round(npix / ms_graph / 1000.0, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_graph, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 3M-Gaussian scene in a 40^3 box, random dL/dimage)",
        "config": {"workload": f"batched fwd+bwd of {args.views} orbit cameras (radius 8-30, "
                               f"resolutions {'/'.join(res)}) over {P} Gaussians, SH3, "
                               "wodilate+OGFS defaults, one launch per stage, graph replay",
                   "gaussians": P, "views": args.views, "pixels": npix, "kv_pairs": K,
                   "visible_pairs": Pv, "work_shape": list(_lib.get_work_shape())},
        "batched": {"graph_ms": round(ms_graph, 3), "eager_ms": round(ms_eager, 3),
                    "ms_per_view": round(ms_graph / args.views, 3),
                    "autotune_ms": {k: round(v, 3) for k, v in (tuned or {}).items()},
                    "blend_config": getattr(br, "tuned", "launch_xcd32"),
                    "stages": stb, "hbm_stages": hbm_frac(stb, nbb)},
        "sequential_plain": {"ms": round(ms_seq, 3), "Mpix/s": round(npix / ms_seq / 1000.0, 3),
                             "batched_speedup": round(ms_seq / ms_graph, 3)},
        "per_view_plain_1080p": {"ms_per_view": round(ms1, 3),
                                 "Mpix/s": round(1920 * 1080 / ms1 / 1000.0, 3),
                                 "visible": Pv1, "kv_pairs": K1, "stages": st1,
                                 "hbm_stages": hbm_frac(st1, nb1)},
        "roofline": {"kernel": dominant, "launch_ms": kern["ms"], "bound": "hbm",
                     "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
    }
    g_t = (means, scales, rots, opac, shs)
    br = graph = grads = None
    torch.cuda.empty_cache()
    if not args.no_1080p_batch:
        line["batched_16x1080p"] = batch_1080p_leg(g_t, args.views, args.steps, args.warmup, plain_step)
        torch.cuda.empty_cache()
    if not args.no_train:
        line["train_step_1080p"] = train_leg(g_t, cam1, max(3, args.steps // 2), args.warmup)
    if not args.no_lod:
        del means, scales, rots, opac, shs, g_t
        torch.cuda.empty_cache()
        line["lod_c4"] = lod_leg(max(2, args.steps // 2), args.warmup)
    if not args.no_cpu_baseline:
        cb, acc = cpu_baseline(g, cam1, dL1, args.cpu_band, gpu_band)
        line["accuracy"] = acc
        if not args.no_cpu_configs:
            cb["configs"] = cpu_configs(cb["cores"])
        line["cpu_baseline"] = cb
    print(json.dumps(line))


if __name__ == "__main__":
    main()
