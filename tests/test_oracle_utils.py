"""Pins of the oracle's per-point utilities against closed forms computed in
FP64 numpy: frustum flags (visible_flag_kernel.cu:19-81, rasterizer_impl.cu:97-109)
and the two 3-sigma radius rules (forward.cu:187-263 with no dilation,
compute_radius_kernel_v2.cu:180-229 with the 0.3 floor)."""
import math

import numpy as np
import pytest

from filtergs_amd.camera import make_camera


@pytest.fixture(scope="module")
def cam():
    return make_camera(320, 240, center=(0.0, -10.0, 0.5), fovx_deg=60.0)


def _hom(p, m):
    """[p, 1] times the stored (transposed, row-vector) 4x4 matrix, in FP64."""
    return np.c_[p.astype(np.float64), np.ones(len(p))] @ np.asarray(m, np.float64)


def test_visible_flag_closed_form(orc, cam):
    rng = np.random.default_rng(3)
    xyz = rng.uniform(-12, 12, size=(4000, 3)).astype(np.float32)
    pad = 0.1
    valid, depth, pp = orc.visible_flag(xyz, cam.full_proj_transform, pad)
    h = _hom(xyz, cam.full_proj_transform)
    w = h[:, 3]
    proj = h[:, :3] / w[:, None]
    np.testing.assert_allclose(depth, w, rtol=1e-5, atol=1e-5)
    ok = np.abs(w) > 1e-2
    np.testing.assert_allclose(pp[ok], proj[ok], rtol=2e-4, atol=2e-4)
    expect = (w > 0) & (w < 100) & (np.abs(proj[:, 0]) < 1 + pad) & (np.abs(proj[:, 1]) < 1 + pad)
    # away from every threshold the flag is fixed by the closed form
    margin = np.minimum.reduce([np.abs(w), np.abs(w - 100),
                                np.abs(np.abs(proj[:, 0]) - (1 + pad)),
                                np.abs(np.abs(proj[:, 1]) - (1 + pad))])
    far = margin > 1e-3
    assert far.sum() > 3500
    assert np.array_equal(valid[far], expect[far])
    assert 200 < expect.sum() < 3800


def test_mark_visible_is_view_depth_test(orc, cam):
    rng = np.random.default_rng(4)
    xyz = rng.uniform(-12, 12, size=(4000, 3)).astype(np.float32)
    present = orc.mark_visible(xyz, cam.world_view_transform, cam.full_proj_transform)
    z = _hom(xyz, cam.world_view_transform)[:, 2]
    far = np.abs(z - 0.2) > 1e-4
    assert np.array_equal(present[far], (z > 0.2)[far])
    assert 100 < present.sum() < 3900


def _on_axis(cam, zs):
    """Points on the optical axis at view depths zs (camera at y = -10 looking at the origin)."""
    c = np.asarray(cam.center, np.float64)
    fwd = -c / np.linalg.norm(c)
    return (c[None, :] + zs[:, None] * fwd[None, :]).astype(np.float32)


def _unit_quats(n, seed):
    q = np.random.default_rng(seed).normal(size=(n, 4))
    return (q / np.linalg.norm(q, axis=1, keepdims=True)).astype(np.float32)


def test_radius_aa_isotropic_closed_form(orc, cam):
    """Isotropic Gaussian of scale s at view depth z on the axis: Sigma2D = (f s/z)^2 I
    (no dilation), disc clamp 0.1: r = 3 sqrt((f s/z)^2 + sqrt(0.1)); any unit
    quaternion gives the same radius; outside |ndc| > 1.3 the radius is 0."""
    zs = np.array([2.0, 4.0, 7.0, 9.0])
    s = np.array([0.02, 0.05, 0.3, 0.01])
    means = _on_axis(cam, zs)
    f = cam.width / (2 * cam.tanfovx)
    expect = 3 * np.sqrt((f * s / zs) ** 2 + math.sqrt(0.1))
    for q in (np.tile([1.0, 0, 0, 0], (4, 1)).astype(np.float32), _unit_quats(4, 5)):
        r = orc.radius_aa(means, np.repeat(s[:, None], 3, 1).astype(np.float32), q,
                          cam.world_view_transform, cam.full_proj_transform, cam.tanfovx,
                          cam.tanfovy, cam.height, cam.width)
        np.testing.assert_allclose(r, expect, rtol=2e-5)
    # far off to the side: culled by the 1.3 NDC guard band
    side = means.copy()
    side[:, 0] += 30.0
    r = orc.radius_aa(side, np.full((4, 3), 0.1, np.float32), _unit_quats(4, 6),
                      cam.world_view_transform, cam.full_proj_transform, cam.tanfovx,
                      cam.tanfovy, cam.height, cam.width)
    assert (r == 0).all()


def test_radius_v2_floor_closed_form(orc, cam):
    """compute_radius_cuda: the same sandwich with max(., 0.3) on the diagonal:
    r = 3 sqrt(max((f s/z)^2, 0.3) + sqrt(0.1))."""
    zs = np.array([2.0, 4.0, 7.0, 9.0])
    s = np.array([0.02, 0.05, 0.3, 0.001])
    means = _on_axis(cam, zs)
    f = cam.width / (2 * cam.tanfovx)
    expect = 3 * np.sqrt(np.maximum((f * s / zs) ** 2, 0.3) + math.sqrt(0.1))
    r = orc.radius_v2(means, np.repeat(s[:, None], 3, 1).astype(np.float32), _unit_quats(4, 7),
                      cam.full_proj_transform, cam.world_view_transform, f, f, cam.tanfovx,
                      cam.tanfovy)
    np.testing.assert_allclose(r, expect, rtol=2e-5)
