"""Pins the oracle's integer stage (rect, keys, sort, ranges) by brute force
and hand-worked cases, and its blend by closed forms and an independent
per-pixel compositing written with numpy cumulative products."""
import numpy as np
import pytest

from tests._scenes import orc_settings, small_scene


def test_higher_msb_is_bit_length(orc):
    for n in list(range(1, 5000)) + [2 ** k + d for k in range(12, 21) for d in (-1, 0, 1)]:
        assert orc.higher_msb(n) == n.bit_length(), n


@pytest.mark.parametrize("px,py,r,expect", [
    # hand-worked: xmin = trunc((px-r)/16), xmax = trunc((px+r+15)/16), clamped to [0, grid]
    (100.3, 40.0, 5, (5, 2, 7, 3)),      # (95.3/16=5.96, 35/16=2.19, 120.3/16=7.52, 60/16=3.75)
    (16.0, 16.0, 1, (0, 0, 2, 2)),       # (15/16, 15/16, 32/16=2, 2)
    (-30.0, 10.0, 3, (0, 0, 0, 1)),      # left of the image: empty in x
    (500.0, 300.0, 400, (6, 0, 20, 15)), # clamped to the 20x15 grid
    (31.9, 0.0, 1, (1, 0, 2, 1)),        # 30.9/16=1.93, 47.9/16=2.99 -> [1,2)
])
def test_get_rect_hand_cases(orc, px, py, r, expect):
    assert tuple(orc.get_rect(px, py, r, 20, 15)) == expect


def _brute_bin(radii, rect, depths, gx, gy):
    pairs = []
    seq = 0
    for i in range(len(radii)):
        if radii[i] <= 0:
            continue
        dbits = int(np.float32(depths[i]).view(np.uint32))
        x0, y0, x1, y1 = (int(v) for v in rect[i])
        for y in range(y0, y1):
            for x in range(x0, x1):
                pairs.append((((y * gx + x) << 32) | dbits, seq, i))
                seq += 1
    pairs.sort()
    keys = [p[0] for p in pairs]
    vals = [p[2] for p in pairs]
    ranges = np.zeros((gx * gy, 2), np.int64)
    for k, key in enumerate(keys):
        t = key >> 32
        if k == 0 or (keys[k - 1] >> 32) != t:
            ranges[t, 0] = k
        ranges[t, 1] = k + 1
    return np.array(keys, np.uint64), np.array(vals, np.uint32), ranges


def test_binning_matches_brute_force(orc):
    g, cam = small_scene(n=1500, W=96, H=80, seed=11)
    s = orc_settings(orc, cam, sh_degree=3, variant=0)
    pre = orc.preprocess(s, g["means3D"], g["opacities"], g["scales"], g["rotations"], g["shs"])
    # inject exact depth ties to exercise stability
    pre["depths"][5:40] = pre["depths"][5]
    keys, vals, ranges = orc.bin_gaussians(pre["radii"], pre["rect"], pre["depths"], 96, 80)
    bk, bv, br = _brute_bin(pre["radii"], pre["rect"], pre["depths"], 6, 5)
    assert len(keys) > 1000
    assert np.array_equal(keys, bk)
    assert np.array_equal(vals, bv)
    assert np.array_equal(ranges.astype(np.int64), br)


def test_binning_empty(orc):
    keys, vals, ranges = orc.bin_gaussians(np.zeros(4, np.int32), np.zeros((4, 4), np.int32),
                                           np.ones(4, np.float32), 32, 32)
    assert keys.size == 0 and vals.size == 0 and not ranges.any()


def _one_tile_settings(orc, bg=(0.1, 0.2, 0.3)):
    from filtergs_amd.camera import make_camera
    cam = make_camera(16, 16, center=(0, -4, 0))
    return orc_settings(orc, cam, bg=bg)


def test_blend_single_gaussian_at_pixel_centre(orc):
    s = _one_tile_settings(orc)
    means2D = np.array([[5.0, 7.0]])
    conic = np.array([[0.5, 0.1, 0.4, 0.6]])
    col = np.array([[0.9, 0.5, 0.2]])
    out = orc.blend_fwd(s, 1, np.array([[0, 1]]), np.array([0]), means2D, col, conic, mode="f64")
    bg = np.array([0.1, 0.2, 0.3], np.float32).astype(np.float64)
    np.testing.assert_allclose(out["color"][:, 7, 5], 0.6 * col[0] + 0.4 * bg, rtol=1e-12)
    assert out["point_id"][7, 5] == 0 and out["n_contrib"][7, 5] == 1
    np.testing.assert_allclose(out["point_weight"][0], 0.6)


def test_blend_alpha_clamp_and_empty(orc):
    s = _one_tile_settings(orc)
    out = orc.blend_fwd(s, 1, np.array([[0, 1]]), np.array([0]), np.array([[3.0, 3.0]]),
                        np.array([[1.0, 1.0, 1.0]]), np.array([[1.0, 0.0, 1.0, 1.0]]), mode="f64")
    # opacity 1 is clamped to alpha 0.99 at the centre
    np.testing.assert_allclose(out["final_T"][3, 3], 1 - np.float32(0.99), rtol=1e-7)
    empty = orc.blend_fwd(s, 1, np.array([[0, 0]]), np.array([0]), np.array([[3.0, 3.0]]),
                          np.array([[1.0, 1.0, 1.0]]), np.array([[1.0, 0.0, 1.0, 1.0]]), mode="f64")
    bg = np.array([0.1, 0.2, 0.3], np.float32)
    assert np.allclose(empty["color"], bg[:, None, None])
    assert (empty["final_T"] == 1).all() and (empty["point_id"] == -1).all()


def _composite_numpy(s, ranges, vals, means2D, rgb, co):
    """Per-pixel front-to-back compositing with cumulative products."""
    H, W = s.image_height, s.image_width
    gx = (W + 15) // 16
    bg = np.array(list(s.bg), np.float64)
    img = np.zeros((3, H, W))
    fT = np.zeros((H, W))
    ncon = np.zeros((H, W), np.int64)
    for py in range(H):
        for px in range(W):
            t = (py // 16) * gx + px // 16
            ids = vals[ranges[t, 0]:ranges[t, 1]].astype(np.int64)
            if ids.size == 0:
                img[:, py, px] = bg
                fT[py, px] = 1
                continue
            dx = means2D[ids, 0] - px
            dy = means2D[ids, 1] - py
            c = co[ids]
            power = -0.5 * (c[:, 0] * dx * dx + c[:, 2] * dy * dy) - c[:, 1] * dx * dy
            alpha = np.minimum(np.float32(0.99), c[:, 3] * np.exp(power))
            use = (power <= 0) & (alpha >= np.float32(1.0) / np.float32(255.0))
            a = np.where(use, alpha, 0.0)
            Tafter = np.cumprod(1 - a)
            stop = np.where(use & (Tafter < np.float32(0.0001)))[0]
            if stop.size:
                a[stop[0]:] = 0.0
            Tbefore = np.concatenate([[1.0], np.cumprod(1 - a)[:-1]])
            img[:, py, px] = (rgb[ids] * (a * Tbefore)[:, None]).sum(0) + np.prod(1 - a) * bg
            fT[py, px] = np.prod(1 - a)
            nz = np.nonzero(a)[0]
            ncon[py, px] = nz[-1] + 1 if nz.size else 0
    return img, fT, ncon


def test_blend_matches_numpy_compositing(orc):
    g, cam = small_scene(n=800, W=48, H=40, seed=12, dist=6.0)
    s = orc_settings(orc, cam, sh_degree=3, variant=0, bg=(0.3, 0.1, 0.7))
    f = orc.render(s, g["means3D"], g["opacities"], g["scales"], g["rotations"], g["shs"],
                   mode="f64")
    pre = f["pre"]
    img, fT, ncon = _composite_numpy(s, f["ranges"], f["vals"], pre["means2D"], pre["rgb"],
                                     pre["conic_opacity"])
    np.testing.assert_allclose(f["color"], img, atol=1e-12)
    np.testing.assert_allclose(f["final_T"], fT, atol=1e-12)
    assert np.array_equal(f["n_contrib"].astype(np.int64), ncon)
    assert (f["final_T"] >= np.float32(0.0001)).all()   # early-stop invariant


def test_point_weight_is_max_of_pixel_weights(orc):
    g, cam = small_scene(n=600, W=64, H=48, seed=13, dist=6.0)
    s = orc_settings(orc, cam, sh_degree=3, variant=1)
    f = orc.render(s, g["means3D"], g["opacities"], g["scales"], g["rotations"], g["shs"],
                   mode="f32")
    pid, wp, pw = f["point_id"], f["weight_pixel"], f["point_weight"]
    sel = pid >= 0
    assert sel.any()
    assert (pw[pid[sel]] >= wp[sel]).all()
    assert ((pw > 0) <= np.isin(np.arange(len(pw)), f["vals"])).all()
