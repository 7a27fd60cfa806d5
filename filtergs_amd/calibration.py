"""Scene calibration of the OGFS tile filter from RenderStats.

The flow of the reference's render app, as host code over the statistics the
blend forward reduces on the device:

1. a debug pass renders every frame with statistics on and keeps one record
   per frame (`frame_record`, apps/render.py:465-481);
2. the records are summarised over the frames (`summarize_frames`,
   apps/render.py:518-562): plain means of the per-frame counters, the overall
   used ratio, and the KPC histograms aligned on the union of their bin edges;
3. the scene's redundancy indicators are written once to `scene_lambda.yml`
   and never overwritten (`save_scene_lambda`, apps/render.py:568-589);
4. later passes read `avg_bar_g` back (`load_scene_avg_bar_g`,
   apps/render.py:174-189) and scale the configured lambda_G into the tile
   threshold tau = lambda_G * avg_bar_g (`resolve_ogfs_controls`,
   apps/render.py:636-655);
5. tau becomes the rasterizer's `ogfs_epsilon_scale` / `ogfs_epsilon_max`
   (`lambda_g_to_scale_max`, `apply_lambda_g_override`,
   apps/render.py:24-66).

Everything here is host arithmetic on a few dozen floats per frame; the
per-pair reductions behind each record happen in the blend forward.
"""
from __future__ import annotations

import math
import os

import yaml

EPSILON_PIX = 1.0 / 255.0

SCENE_LAMBDA_FILE = "scene_lambda.yml"

# Keys of scene_lambda.yml, in the order the reference writes them
# (apps/render.py:576-586).
SCENE_LAMBDA_KEYS = (
    "avg_psnr", "avg_ssim", "avg_lpips", "avg_bar_g", "avg_num_kv_pairs",
    "avg_used_kv_pairs", "avg_unused_kv_pairs", "avg_used_kv_ratio",
    "total_used_kv_ratio",
)


def round6(value):
    """None stays None; anything else becomes a float rounded to 6 places
    (apps/render.py:166-169)."""
    return None if value is None else float(round(float(value), 6))


def lambda_g_to_scale_max(lambda_g, block_x=16, block_y=16):
    """Map a tile threshold lambda_G to (ogfs_epsilon_scale, ogfs_epsilon_max)
    (apps/render.py:24-37).

    A tile of block_x*block_y pixels each at the 1/255 floor has energy
    G_raw = 1 - (1 - 1/255)^(block pixels); the scale is the position of the
    (floored) target between 1/255 and G_raw, clamped at 0, and the maximum is
    the floored target itself.
    """
    target = max(float(lambda_g), EPSILON_PIX)
    keep = max(0.0, 1.0 - EPSILON_PIX)
    g_raw = 1.0 - pow(keep, float(block_x * block_y))
    span = g_raw - EPSILON_PIX
    scale = 0.0 if abs(span) < 1e-6 else (target - EPSILON_PIX) / span
    return max(0.0, float(scale)), float(max(target, EPSILON_PIX))


def apply_lambda_g_override(rasterizer, lambda_g, ogfs_energy_floor=None,
                            ogfs_radius_ratio_min=None):
    """Replace the OGFS fields of `rasterizer.raster_settings` in place
    (apps/render.py:40-66).

    Does nothing when every control is None, or when the rasterizer has no
    NamedTuple settings carrying the OGFS epsilon fields.
    """
    if lambda_g is None and ogfs_energy_floor is None and ogfs_radius_ratio_min is None:
        return
    rs = getattr(rasterizer, "raster_settings", None)
    if rs is None or not hasattr(rs, "_replace"):
        return
    fields = set(getattr(rs, "_fields", ()))
    if not {"ogfs_epsilon_scale", "ogfs_epsilon_max"} <= fields:
        return
    changes = {}
    if lambda_g is not None:
        changes["ogfs_epsilon_scale"], changes["ogfs_epsilon_max"] = lambda_g_to_scale_max(lambda_g)
    if ogfs_energy_floor is not None and "ogfs_energy_floor" in fields:
        changes["ogfs_energy_floor"] = float(ogfs_energy_floor)
    if ogfs_radius_ratio_min is not None and "ogfs_radius_ratio_min" in fields:
        changes["ogfs_radius_ratio_min"] = float(ogfs_radius_ratio_min)
    if changes:
        rasterizer.raster_settings = rs._replace(**changes)


def frame_record(latest_stats):
    """Per-frame statistics record from a rasterizer's `latest_stats` dict
    (apps/render.py:465-481).

    Missing counters default as the reference's do: used = all pairs, unused =
    the rest, the used ratio 0 for a frame without pairs.
    """
    st = latest_stats or {}
    kv = float(st.get("num_kv_pairs", 0.0))
    used = float(st.get("used_kv_pairs", kv))
    unused = float(st.get("unused_kv_pairs", max(0.0, kv - used)))
    return {
        "bar_g": float(st.get("bar_g", 0.0)),
        "avg_kv_per_gaussian": float(st.get("avg_kv_per_gaussian", 0.0)),
        "avg_kv_per_tile": float(st.get("avg_kv_per_tile", 0.0)),
        "num_kv_pairs": kv,
        "used_kv_pairs": used,
        "unused_kv_pairs": unused,
        "used_kv_ratio": used / kv if kv > 0 else 0.0,
        "kpc_edges": [float(x) for x in (st.get("kpc_edges") or [])],
        "kpc_hist": [float(x) for x in (st.get("kpc_hist") or [])],
    }


def _mean(values):
    return sum(values) / len(values) if values else 0.0


def _mean_or_none(values):
    return sum(values) / len(values) if values else None


def summarize_frames(frames, psnr=(), ssim=(), lpips=()):
    """Scene summary of per-frame records (apps/render.py:518-562).

    `psnr`, `ssim` and `lpips` are the per-frame image metrics when the caller
    measured them (the reference averages them over the frames that have
    one); they only feed the `avg_*` quality keys.

    KPC histograms: the bin edges of all frames are merged into one sorted
    set; each frame's finite bin values are added at their edge's position and
    counted; `kpc_hist_avg` divides by that count (0 where no frame had a
    value). Without any edge the KPC keys are absent.
    """
    frames = list(frames)
    summary = {
        "num_frames": len(frames),
        "avg_psnr": _mean_or_none(list(psnr)),
        "avg_ssim": _mean_or_none(list(ssim)),
        "avg_lpips": _mean_or_none(list(lpips)),
    }
    kv = [float(f.get("num_kv_pairs", 0.0)) for f in frames]
    used = [float(f.get("used_kv_pairs", 0.0)) for f in frames]
    summary["avg_bar_g"] = _mean([float(f.get("bar_g", 0.0)) for f in frames])
    summary["avg_num_kv_pairs"] = _mean(kv)
    summary["avg_used_kv_pairs"] = _mean(used)
    summary["avg_unused_kv_pairs"] = _mean([float(f.get("unused_kv_pairs", 0.0)) for f in frames])
    summary["avg_used_kv_ratio"] = _mean([float(f.get("used_kv_ratio", 0.0)) for f in frames])
    kv_sum = sum(kv)
    summary["total_used_kv_ratio"] = sum(used) / kv_sum if kv_sum > 0 else 0.0

    edges = sorted({float(e) for f in frames for e in f.get("kpc_edges", []) if e is not None})
    if edges:
        slot = {e: i for i, e in enumerate(edges)}
        hsum = [0.0] * len(edges)
        hcnt = [0] * len(edges)
        for f in frames:
            for e, v in zip(f.get("kpc_edges", []), f.get("kpc_hist", [])):
                if e is None:
                    continue
                v = float(v)
                if math.isfinite(v):
                    i = slot[float(e)]
                    hsum[i] += v
                    hcnt[i] += 1
        summary["kpc_bin_edges"] = edges
        summary["kpc_hist_sum"] = hsum
        summary["kpc_hist_avg"] = [s / c if c > 0 else 0.0 for s, c in zip(hsum, hcnt)]
    return summary


def scene_lambda(summary):
    """The scene_lambda.yml record: the summary's indicators rounded to 6
    places, None where absent (apps/render.py:576-586)."""
    return {k: round6(summary.get(k)) for k in SCENE_LAMBDA_KEYS}


def save_scene_lambda(scene_dir, summary):
    """Write `scene_dir/scene_lambda.yml` unless it exists; the file is
    immutable once created (apps/render.py:568-589). Returns True when it
    wrote the file."""
    path = os.path.join(scene_dir, SCENE_LAMBDA_FILE)
    if os.path.exists(path):
        return False
    with open(path, "w") as f:
        yaml.safe_dump(scene_lambda(summary), f, sort_keys=False)
    return True


def load_scene_avg_bar_g(scene_dir):
    """`avg_bar_g` from `scene_dir/scene_lambda.yml`, or None when the file is
    missing, unreadable, lacks the key or holds a non-finite value
    (apps/render.py:174-189)."""
    path = os.path.join(scene_dir, SCENE_LAMBDA_FILE)
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            data = yaml.safe_load(f) or {}
        v = data.get("avg_bar_g")
        if v is None:
            return None
        v = float(v)
        return v if math.isfinite(v) else None
    except Exception as exc:  # the reference reports and carries on
        print(f"[calibration] failed to read {path}: {exc}")
        return None


def resolve_ogfs_controls(cfg, scene_dir):
    """Effective OGFS controls of a render config (apps/render.py:636-655).

    `cfg` is the render config mapping (`ogfs_lambda_G`, or `lambda_G`;
    `ogfs_energy_floor`; `ogfs_radius_ratio_min`). With a lambda_G and a
    calibrated scene, lambda_G is scaled by the scene's `avg_bar_g`; with a
    lambda_G, an unset energy floor defaults to 0.01 and an unset radius
    ratio to 0.1. Returns a dict with `lambda_g` (effective), `lambda_g_base`,
    `avg_bar_g`, `energy_floor` and `radius_ratio_min`.
    """
    lam = cfg.get("ogfs_lambda_G", cfg.get("lambda_G", None))
    floor = cfg.get("ogfs_energy_floor", None)
    rmin = cfg.get("ogfs_radius_ratio_min", None)
    base = lam
    bar_g = load_scene_avg_bar_g(scene_dir)
    if lam is not None and bar_g is not None:
        lam = float(lam) * float(bar_g)
    if lam is not None and floor is None:
        floor = 0.01
    if lam is not None and rmin is None:
        rmin = 0.1
    return {"lambda_g": lam, "lambda_g_base": base, "avg_bar_g": bar_g,
            "energy_floor": floor, "radius_ratio_min": rmin}
