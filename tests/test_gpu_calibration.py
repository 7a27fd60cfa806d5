"""Scene calibration end to end on the device: RenderStats of real renders ->
frame records -> scene summary -> scene_lambda.yml -> effective OGFS controls
-> rasterizer settings (apps/render.py:24-66, 465-481, 518-589, 636-655).
The per-frame bar_g is checked against the oracle's per-pair sums, and the
render under the calibrated settings against the oracle."""
import numpy as np
import pytest
import torch

from filtergs_amd import calibration as cal
from filtergs_amd.camera import make_camera
from filtergs_amd.rasterizer import GaussianRasterizationSettings, GaussianRasterizer
from tests._scenes import orc_settings, small_scene
from tests.test_gpu_parity import _check_forward, _t, gpu_forward

pytestmark = pytest.mark.gpu


def _module(cam, bg=(0.1, 0.2, 0.3)):
    rs = GaussianRasterizationSettings(
        image_height=cam.height, image_width=cam.width, tanfovx=cam.tanfovx, tanfovy=cam.tanfovy,
        bg=_t(bg), scale_modifier=1.0, viewmatrix=_t(cam.world_view_transform),
        projmatrix=_t(cam.full_proj_transform), sh_degree=3, campos=_t(cam.center),
        prefiltered=False, debug=False)
    return GaussianRasterizer(rs)


def _render(r, g):
    P = g["means3D"].shape[0]
    with torch.no_grad():
        return r(means3D=_t(g["means3D"]), means2D=torch.zeros(P, 3, device="cuda"),
                 opacities=_t(g["opacities"].reshape(P, 1)), shs=_t(g["shs"]),
                 scales=_t(g["scales"]), rotations=_t(g["rotations"]))


def test_calibration_flow(orc, tmp_path):
    g, _ = small_scene(n=3000, W=256, H=192, seed=171, dist=6.0)
    cams = [make_camera(256, 192, center=c, fovx_deg=60.0)
            for c in ((0.0, -6.0, 0.5), (4.0, -4.5, 1.5), (-3.0, -5.0, -1.0))]
    frames = []
    for cam in cams:
        r = _module(cam)
        _render(r, g)
        rec = cal.frame_record(r.latest_stats)
        frames.append(rec)
        s = orc_settings(orc, cam, sh_degree=3, variant=0)
        f = orc.render(s, g["means3D"], g["opacities"], g["scales"], g["rotations"], g["shs"],
                       mode="f32", stats=True)
        kv, rg = f["kv_alpha_sum"].astype(np.float64), f["ranges"]
        bar_g = np.mean([kv[a:e].mean() if e > a else 0.0 for a, e in rg])
        assert abs(rec["bar_g"] - bar_g) <= 1e-5
        assert rec["num_kv_pairs"] == len(kv) > 0
        assert rec["used_kv_pairs"] + rec["unused_kv_pairs"] <= rec["num_kv_pairs"]
        assert len(rec["kpc_edges"]) == len(rec["kpc_hist"]) == 11
    summary = cal.summarize_frames(frames)
    assert summary["avg_bar_g"] == pytest.approx(np.mean([f["bar_g"] for f in frames]), rel=1e-12)
    assert summary["kpc_bin_edges"] == sorted(frames[0]["kpc_edges"])
    assert cal.save_scene_lambda(str(tmp_path), summary)

    ctl = cal.resolve_ogfs_controls({"ogfs_lambda_G": 2.0}, str(tmp_path))
    assert ctl["lambda_g"] == pytest.approx(2.0 * cal.round6(summary["avg_bar_g"]), rel=1e-12)
    r = _module(cams[1])
    cal.apply_lambda_g_override(r, ctl["lambda_g"], ctl["energy_floor"], ctl["radius_ratio_min"])
    rs = r.raster_settings
    assert (rs.ogfs_energy_floor, rs.ogfs_radius_ratio_min) == (0.01, 0.1)
    color, radii = _render(r, g)[:2]
    ogfs = dict(ogfs_epsilon_scale=rs.ogfs_epsilon_scale, ogfs_epsilon_max=rs.ogfs_epsilon_max,
                ogfs_energy_floor=rs.ogfs_energy_floor,
                ogfs_radius_ratio_min=rs.ogfs_radius_ratio_min)
    out = gpu_forward(g, cams[1], bg=(0.1, 0.2, 0.3), **ogfs)
    assert torch.equal(out[5], radii) and torch.equal(out[1], color)
    # the calibrated settings move the OGFS radii away from the default ones
    base = gpu_forward(g, cams[1], bg=(0.1, 0.2, 0.3))
    assert not torch.equal(base[5], radii)
    _check_forward(orc, g, cams[1], 0, **ogfs)
