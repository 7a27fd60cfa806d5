"""Scene calibration flow of the render app (apps/render.py:24-66, 166-189,
465-481, 518-589, 636-655): host arithmetic, checked against closed forms and
hand-worked frames."""
import math

import pytest
import torch
import yaml

from filtergs_amd import calibration as cal
from filtergs_amd.rasterizer import GaussianRasterizationSettings, parse_stats

EPS = 1.0 / 255.0
# Energy of a 16x16 tile of pixels at the 1/255 floor, computed through log1p
# rather than pow so a slip in the exponent or the base shows.
G_RAW = -math.expm1(256.0 * math.log1p(-EPS))


def test_lambda_g_to_scale_max_closed_form():
    assert G_RAW == pytest.approx(0.634283, abs=1e-6)
    s, m = cal.lambda_g_to_scale_max(EPS)
    assert s == 0.0 and m == pytest.approx(EPS, rel=1e-15)
    s, m = cal.lambda_g_to_scale_max(G_RAW)
    assert s == pytest.approx(1.0, rel=1e-12) and m == G_RAW
    s, m = cal.lambda_g_to_scale_max(0.5 * (EPS + G_RAW))
    assert s == pytest.approx(0.5, rel=1e-12)
    # below the floor: clamped to the floor, scale 0
    assert cal.lambda_g_to_scale_max(0.0) == (0.0, EPS)
    assert cal.lambda_g_to_scale_max(-3.0) == (0.0, EPS)
    # a 1-pixel block has G_raw = 1/255: degenerate span, scale 0
    assert cal.lambda_g_to_scale_max(0.3, 1, 1) == (0.0, 0.3)
    # block size enters through the exponent
    g4 = 1.0 - (1.0 - EPS) ** 16
    s, _ = cal.lambda_g_to_scale_max(g4, 4, 4)
    assert s == pytest.approx(1.0, rel=1e-12)


def _settings():
    return GaussianRasterizationSettings(
        image_height=8, image_width=8, tanfovx=0.5, tanfovy=0.5, bg=torch.zeros(3),
        scale_modifier=1.0, viewmatrix=torch.eye(4), projmatrix=torch.eye(4), sh_degree=0,
        campos=torch.zeros(3), prefiltered=False, debug=False)


class _Holder:
    def __init__(self, rs):
        self.raster_settings = rs


def test_apply_lambda_g_override():
    h = _Holder(_settings())
    before = h.raster_settings
    cal.apply_lambda_g_override(h, None)
    assert h.raster_settings is before
    cal.apply_lambda_g_override(h, 0.2, ogfs_energy_floor=0.01, ogfs_radius_ratio_min=0.1)
    rs = h.raster_settings
    want_s, want_m = cal.lambda_g_to_scale_max(0.2)
    assert (rs.ogfs_epsilon_scale, rs.ogfs_epsilon_max) == (want_s, want_m)
    assert (rs.ogfs_energy_floor, rs.ogfs_radius_ratio_min) == (0.01, 0.1)
    assert rs.image_width == 8 and rs.ogfs_enable_stats is True
    # only the energy floor
    h2 = _Holder(_settings())
    cal.apply_lambda_g_override(h2, None, ogfs_energy_floor=0.3)
    assert h2.raster_settings.ogfs_energy_floor == 0.3
    assert h2.raster_settings.ogfs_epsilon_scale == _settings().ogfs_epsilon_scale
    # settings without the OGFS fields or without _replace are left alone
    class Plain:
        raster_settings = {"ogfs_epsilon_scale": 1.0}
    p = Plain()
    cal.apply_lambda_g_override(p, 0.2)
    assert p.raster_settings == {"ogfs_epsilon_scale": 1.0}


def test_frame_record_defaults():
    r = cal.frame_record(None)
    assert r["num_kv_pairs"] == 0.0 and r["used_kv_ratio"] == 0.0
    assert r["kpc_edges"] == [] and r["kpc_hist"] == []
    r = cal.frame_record({"num_kv_pairs": 40})
    assert (r["used_kv_pairs"], r["unused_kv_pairs"], r["used_kv_ratio"]) == (40.0, 0.0, 1.0)
    r = cal.frame_record({"num_kv_pairs": 40, "used_kv_pairs": 10})
    assert (r["unused_kv_pairs"], r["used_kv_ratio"]) == (30.0, 0.25)
    r = cal.frame_record({"num_kv_pairs": 40, "used_kv_pairs": 10, "unused_kv_pairs": 7})
    assert r["unused_kv_pairs"] == 7.0


def _packed_stats(T=2):
    """A packed RenderStats tensor in the rasterizer's layout: 9 head values,
    K KPC edges, K KPC values, T tile ratios, T tile totals (K = 3 here)."""
    head = [0.25, 3.0, 12.0, 24.0, float(T), 8.0, 18.0, 6.0, 3.0]
    edges = [0.0, 0.5, 1.0]
    hist = [2.0, float("nan"), 4.0]
    return torch.tensor(head + edges + hist + [0.5] * T + [12.0] * T, dtype=torch.float64)


def test_frame_record_from_packed_stats():
    r = cal.frame_record(parse_stats(_packed_stats()))
    assert r["bar_g"] == 0.25 and r["num_kv_pairs"] == 24.0
    assert (r["used_kv_pairs"], r["unused_kv_pairs"], r["used_kv_ratio"]) == (18.0, 6.0, 0.75)
    assert r["kpc_edges"] == [0.0, 0.5, 1.0]
    assert r["kpc_hist"][0] == 2.0 and math.isnan(r["kpc_hist"][1])


def test_summarize_frames_by_hand():
    f1 = {"bar_g": 0.2, "num_kv_pairs": 100.0, "used_kv_pairs": 60.0, "unused_kv_pairs": 40.0,
          "used_kv_ratio": 0.6, "kpc_edges": [0.0, 1.0, 2.0], "kpc_hist": [1.0, float("nan"), 3.0]}
    f2 = {"bar_g": 0.4, "num_kv_pairs": 300.0, "used_kv_pairs": 60.0, "unused_kv_pairs": 240.0,
          "used_kv_ratio": 0.2, "kpc_edges": [1.0, 2.0, 5.0], "kpc_hist": [7.0, 5.0, float("inf")]}
    f3 = {"bar_g": 0.0, "num_kv_pairs": 0.0, "used_kv_pairs": 0.0, "unused_kv_pairs": 0.0,
          "used_kv_ratio": 0.0, "kpc_edges": [], "kpc_hist": []}
    s = cal.summarize_frames([f1, f2, f3], psnr=[30.0, 32.0])
    assert s["num_frames"] == 3
    assert s["avg_bar_g"] == pytest.approx(0.2)
    assert s["avg_num_kv_pairs"] == pytest.approx(400.0 / 3)
    assert s["avg_used_kv_pairs"] == pytest.approx(40.0)
    assert s["avg_unused_kv_pairs"] == pytest.approx(280.0 / 3)
    assert s["avg_used_kv_ratio"] == pytest.approx(0.8 / 3)
    # overall ratio weights frames by their pairs: 120 / 400, not the mean ratio
    assert s["total_used_kv_ratio"] == pytest.approx(0.3)
    assert s["avg_psnr"] == 31.0 and s["avg_ssim"] is None
    assert s["kpc_bin_edges"] == [0.0, 1.0, 2.0, 5.0]
    # edge 1.0: NaN of f1 skipped, 7 of f2; edge 2.0: 3 + 5; edge 5.0: inf skipped
    assert s["kpc_hist_sum"] == [1.0, 7.0, 8.0, 0.0]
    assert s["kpc_hist_avg"] == [1.0, 7.0, 4.0, 0.0]


def test_summarize_no_frames_and_no_edges():
    s = cal.summarize_frames([])
    assert s["avg_bar_g"] == 0.0 and s["total_used_kv_ratio"] == 0.0
    assert "kpc_bin_edges" not in s


def test_scene_lambda_file_is_write_once(tmp_path):
    s = {"avg_bar_g": 0.123456789, "avg_num_kv_pairs": 10.0, "total_used_kv_ratio": 1 / 3}
    assert cal.save_scene_lambda(str(tmp_path), s)
    data = yaml.safe_load(open(tmp_path / cal.SCENE_LAMBDA_FILE))
    assert list(data) == list(cal.SCENE_LAMBDA_KEYS)
    assert data["avg_bar_g"] == 0.123457 and data["total_used_kv_ratio"] == 0.333333
    assert data["avg_psnr"] is None
    assert not cal.save_scene_lambda(str(tmp_path), {"avg_bar_g": 9.0})
    assert cal.load_scene_avg_bar_g(str(tmp_path)) == 0.123457


def test_load_scene_avg_bar_g_rejects(tmp_path):
    assert cal.load_scene_avg_bar_g(str(tmp_path)) is None
    p = tmp_path / cal.SCENE_LAMBDA_FILE
    p.write_text("avg_bar_g: .nan\n")
    assert cal.load_scene_avg_bar_g(str(tmp_path)) is None
    p.write_text("avg_psnr: 3\n")
    assert cal.load_scene_avg_bar_g(str(tmp_path)) is None
    p.write_text("avg_bar_g: [unclosed\n")
    assert cal.load_scene_avg_bar_g(str(tmp_path)) is None
    p.write_text("")
    assert cal.load_scene_avg_bar_g(str(tmp_path)) is None


def test_resolve_ogfs_controls(tmp_path):
    d = str(tmp_path)
    c = cal.resolve_ogfs_controls({}, d)
    assert c["lambda_g"] is None and c["energy_floor"] is None and c["radius_ratio_min"] is None
    c = cal.resolve_ogfs_controls({"lambda_G": 0.5}, d)
    assert (c["lambda_g"], c["energy_floor"], c["radius_ratio_min"]) == (0.5, 0.01, 0.1)
    cal.save_scene_lambda(d, {"avg_bar_g": 0.25})
    c = cal.resolve_ogfs_controls({"ogfs_lambda_G": 0.5, "lambda_G": 9.0,
                                   "ogfs_radius_ratio_min": 0.7}, d)
    assert c["lambda_g"] == 0.125 and c["lambda_g_base"] == 0.5 and c["avg_bar_g"] == 0.25
    assert (c["energy_floor"], c["radius_ratio_min"]) == (0.01, 0.7)
    c = cal.resolve_ogfs_controls({"ogfs_energy_floor": 0.2}, d)
    assert c["lambda_g"] is None and c["energy_floor"] == 0.2 and c["radius_ratio_min"] is None
