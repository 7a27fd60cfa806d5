"""Benchmark: fwd+bwd of the rasterizer for 3M Gaussians at 1920x1080 on one MI355X.

One step = one training-style pass of one view: forward (preprocess, depth
sort, duplicate, tile sort, ranges, blend) + backward (blend backward,
preprocess backward) through the C ABI, for a fixed seeded synthetic scene
(3M Gaussians in a 40^3 box, SH degree 3, wodilate + OGFS defaults) and a
random dL/dimage.  Inputs are resident in HBM before timing starts.

Prints one JSON line (value = Mpix/s of fwd+bwd, higher is better) with a
roofline entry for the dominant kernel (HIP events recorded on the launch
stream around every kernel of the timed steps) and a CPU baseline (the oracle
on a bounded sample of the same view, timed on this host).
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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--gaussians", type=int, default=3_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-band", type=int, default=256, help="rows of the CPU baseline sample")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after timing, run one device-to-device copy of this many MiB (a "
                         "reference kernel of known traffic for hardware byte counters)")
    return ap.parse_args()


def stage_bytes(P, Pv, K, T, npix, M, passes_tile):
    """Algorithmic HBM bytes per launch of each stage (per view)."""
    return {
        "preprocess": P * (44 + 12 * M) + P * 8 + Pv * 40,
        "depth_sort": 4 * 24 * P,
        "tile_offsets": 8 * P,
        "duplicate": 20 * Pv + 8 * K,
        "tile_sort": passes_tile * 24 * K,
        "ranges": 8 * K + 8 * T,
        "blend_fwd": 40 * K + 28 * npix + 4 * P,
        "blend_bwd": 40 * K + 20 * npix + 36 * Pv,
        "preprocess_bwd": Pv * (44 + 12 * M + 36) + P * (92 + 12 * M),
    }


def cpu_baseline(g, cam, dL, band):
    """Oracle (FP32 emulation) fwd+bwd on a horizontal band of the same view:
    all Gaussians are preprocessed, pixels of the central `band` rows are
    blended forward and backward.  Mpix/s = band pixels / wall time."""
    from filtergs_amd.camera import Camera
    from oracle import pyoracle as orc
    y0 = (cam.height - band) // 2
    sub = Camera(cam.width, band, cam.fx, cam.fy, cam.cx, cam.cy - y0, cam.R, cam.T, cam.znear,
                 cam.zfar)
    # same rays: the band keeps fy and shifts the principal point; tanfovy follows H
    s = orc.settings(band, cam.width, sub.tanfovx, sub.tanfovy, sub.world_view_transform,
                     sub.full_proj_transform, sub.center, sh_degree=3)
    cores = min(os.cpu_count() or 1, 16)
    t0 = time.perf_counter()
    f = orc.render(s, g["means3D"], g["opacities"], g["scales"], g["rotations"], g["shs"],
                   mode="f32", nthreads=cores)
    orc.backward(s, f, np.ascontiguousarray(dL[:, y0:y0 + band, :]), g["means3D"],
                 g["opacities"], g["scales"], g["rotations"], g["shs"], mode="f32")
    dt = time.perf_counter() - t0
    return {"value": round(cam.width * band / dt / 1e6, 4), "unit": "Mpix/s", "cores": cores,
            "kind": "port",
            "sample": f"oracle FP32 fwd+bwd of rows {y0}..{y0 + band} of the {cam.width}x{cam.height} "
                      f"view, all {g['means3D'].shape[0]} Gaussians preprocessed "
                      f"(preprocess/blend-forward on {cores} threads, backward 1 thread), {dt:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import torch

    from filtergs_amd import _lib, ext, workloads
    W, H, P = args.width, args.height, args.gaussians
    g = workloads.gaussians(P, 20.0, 3, seed=0)
    cam = workloads.camera_c1(W, H)
    rng = np.random.default_rng(1)
    dL = rng.normal(size=(3, H, W)).astype(np.float32)
    dev = "cuda"

    def t(a, dt=torch.float32):
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=dev)

    means, scales, rots = t(g["means3D"]), t(g["scales"]), t(g["rotations"])
    opac, shs = t(g["opacities"].reshape(-1, 1)), t(g["shs"])
    vm, pm, cp = t(cam.world_view_transform), t(cam.full_proj_transform), t(cam.center)
    bg = t([0.0, 0.0, 0.0])
    dLt = t(dL)
    empty = torch.Tensor([])

    def step():
        out = ext.rasterize_gaussians(bg, means, empty, opac, scales, rots, 1.0, empty, vm, pm,
                                      cam.tanfovx, cam.tanfovy, H, W, shs, 3, cp, False, True,
                                      False, ogfs_enable_stats=False)
        K, color, pid, wpix, pw, radii, geom, binning, img, _ = out
        ext.rasterize_gaussians_backward(bg, means, radii, empty, scales, rots, 1.0, empty, vm,
                                         pm, cam.tanfovx, cam.tanfovy, dLt, shs, 3, cp, geom, K,
                                         binning, img, False)
        return K, radii

    for _ in range(args.warmup):
        K, radii = step()
    torch.cuda.synchronize()
    Pv = int((radii > 0).sum().item())
    _lib.profile_enable(True)
    _lib.profile_read(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        K, radii = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    prof = _lib.profile_read(reset=True)
    _lib.profile_enable(False)
    if args.copy_probe_mib > 0:
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    ms = elapsed * 1000.0 / args.steps
    npix = W * H
    T = ((W + 15) // 16) * ((H + 15) // 16)
    passes_tile = (T.bit_length() + 7) // 8
    nbytes = stage_bytes(P, Pv, K, T, npix, 16, passes_tile)
    stages = {}
    for name, (tot, calls) in prof.items():
        if calls:
            avg = tot / calls
            stages[name] = {"ms": round(avg, 4), "GB/s": round(nbytes[name] / avg / 1e6, 1)}
    dominant = max(stages, key=lambda k: stages[k]["ms"])
    avg_ms = stages[dominant]["ms"]
    achieved = nbytes[dominant] / avg_ms / 1e6
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dominant}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    line = {
        "metric": "fwd+bwd ms/view & Mpix/s, 3M Gaussians @1080p; batched vs per-view; % HBM peak",
        "value": round(npix / ms / 1000.0, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 3M-Gaussian scene, random dL/dimage)",
        "config": {"workload": f"per-view fwd+bwd, {P} Gaussians in a 40^3 box, {W}x{H}, SH3, "
                               "wodilate+OGFS defaults, stats off",
                   "gaussians": P, "width": W, "height": H, "visible": Pv, "kv_pairs": K,
                   "path": "plain per-view"},
        "roofline": {"kernel": dominant, "bound": "hbm", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "stages": stages,
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(g, cam, dL, args.cpu_band)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
