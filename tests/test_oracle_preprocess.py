"""Pins the oracle's preprocess with checks that do not restate its formulas:
library rotations (scipy), numerically differentiated pinhole projection,
closed forms of the pixel mapping and of the OGFS radius, and the SH
addition theorem / orthonormality."""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from filtergs_amd.camera import make_camera
from tests._scenes import orc_settings


def _front_points(n, rng, cam_dist=6.0):
    return np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)], 1)


@pytest.fixture
def cam():
    return make_camera(320, 240, center=(0.3, -6.0, 0.4), target=(0.0, 0.0, 0.0), fovx_deg=55.0)


def test_cov3d_matches_scipy_rotation(orc, cam):
    rng = np.random.default_rng(1)
    n = 300
    means = _front_points(n, rng)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    scales = np.exp(rng.normal(-2, 0.5, size=(n, 3)))
    s = orc_settings(orc, cam, sh_degree=0, variant=1)
    out = orc.preprocess(s, means, np.full(n, 0.5), scales, q, colors=np.full((n, 3), 0.5),
                         mode="f64")
    assert (out["radii"] > 0).all()
    Rm = Rotation.from_quat(np.concatenate([q[:, 1:], q[:, :1]], axis=1)).as_matrix()
    Sig = Rm @ (scales[:, :, None] ** 2 * np.eye(3)[None]) @ np.transpose(Rm, (0, 2, 1))
    ref = np.stack([Sig[:, 0, 0], Sig[:, 0, 1], Sig[:, 0, 2], Sig[:, 1, 1], Sig[:, 1, 2],
                    Sig[:, 2, 2]], 1)
    np.testing.assert_allclose(out["cov3D"], ref, rtol=1e-10, atol=1e-14)


def _pixel_of(cam, x):
    Xc = cam.R @ x + cam.T
    return np.array([cam.fx * Xc[0] / Xc[2] + cam.cx - 0.5, cam.fy * Xc[1] / Xc[2] + cam.cy - 0.5])


def test_means2d_is_pinhole_projection(orc, cam):
    rng = np.random.default_rng(2)
    n = 200
    means = _front_points(n, rng)
    s = orc_settings(orc, cam, sh_degree=0, variant=1)
    out = orc.preprocess(s, means, np.full(n, 0.5), np.full((n, 3), 0.05),
                         np.tile([1.0, 0, 0, 0], (n, 1)), colors=np.full((n, 3), 0.5), mode="f64")
    ref = np.stack([_pixel_of(cam, m) for m in means])
    # matrices are float32 (as the reference's prepare_camera rounds them)
    np.testing.assert_allclose(out["means2D"], ref, atol=2e-3)
    np.testing.assert_allclose(out["depths"], [(cam.R @ m + cam.T)[2] for m in means], rtol=1e-6)


def test_sigma2d_is_projection_jacobian_sandwich(orc, cam):
    """EWA: Sigma2D = J Sigma3D J^T, J = d pixel / d world at the mean, taken
    here by central differences of the pinhole mapping."""
    rng = np.random.default_rng(3)
    n = 100
    means = _front_points(n, rng) * 0.5
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    scales = np.exp(rng.normal(-1.5, 0.4, size=(n, 3)))
    s = orc_settings(orc, cam, sh_degree=0, variant=0, use_filter=False)
    out = orc.preprocess(s, means, np.full(n, 0.5), scales, q, colors=np.full((n, 3), 0.5),
                         mode="f64")
    assert (out["radii"] > 0).all()
    for i in range(n):
        h = 1e-5
        J = np.zeros((2, 3))
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            J[:, k] = (_pixel_of(cam, means[i] + e) - _pixel_of(cam, means[i] - e)) / (2 * h)
        c = out["cov3D"][i]
        S3 = np.array([[c[0], c[1], c[2]], [c[1], c[3], c[4]], [c[2], c[4], c[5]]])
        S2 = J @ S3 @ J.T
        a, b, cc, _o = out["conic_opacity"][i]
        conic = np.array([[a, b], [b, cc]])
        # conic is the inverse of the (undilated) 2D covariance
        np.testing.assert_allclose(np.linalg.inv(conic), S2, rtol=2e-4, atol=1e-6)


def test_dilation_modes(orc, cam):
    """wodilate+filter: diagonal floor at 0.3; dilate: +0.3 on the diagonal."""
    rng = np.random.default_rng(4)
    n = 200
    means = _front_points(n, rng)
    scales = np.exp(rng.normal(-5, 1.0, size=(n, 3)))   # sub-pixel: the floor is active
    q = np.tile([1.0, 0, 0, 0], (n, 1))
    base = orc.preprocess(orc_settings(orc, cam, variant=0, use_filter=False), means,
                          np.full(n, 0.5), scales, q, colors=np.full((n, 3), 0.5), mode="f64")
    filt = orc.preprocess(orc_settings(orc, cam, variant=0, use_filter=True), means,
                          np.full(n, 0.5), scales, q, colors=np.full((n, 3), 0.5), mode="f64")
    dil = orc.preprocess(orc_settings(orc, cam, variant=1), means, np.full(n, 0.5), scales, q,
                         colors=np.full((n, 3), 0.5), mode="f64")

    def cov2(o):
        a, b, c, _ = o.T
        det = a * c - b * b
        return np.stack([c / det, -b / det, a / det], 1)   # inverse of the conic

    ok = (base["radii"] > 0) & (filt["radii"] > 0) & (dil["radii"] > 0)
    cb, cf, cd = cov2(base["conic_opacity"][ok]), cov2(filt["conic_opacity"][ok]), cov2(dil["conic_opacity"][ok])
    np.testing.assert_allclose(cf[:, 0], np.maximum(cb[:, 0], np.float32(0.3)), rtol=1e-7)
    np.testing.assert_allclose(cf[:, 2], np.maximum(cb[:, 2], np.float32(0.3)), rtol=1e-7)
    np.testing.assert_allclose(cd[:, 0], cb[:, 0] + 0.3, rtol=1e-7)
    np.testing.assert_allclose(cd[:, 1], cb[:, 1], rtol=1e-7, atol=1e-12)


def _sigma_px(conic_opacity):
    a, b, c, _ = conic_opacity.T
    det = a * c - b * b
    S = np.stack([np.stack([c / det, -b / det], -1), np.stack([-b / det, a / det], -1)], -2)
    return np.sqrt(np.linalg.eigvalsh(S)[:, -1])


def test_ogfs_radius_closed_form_defaults(orc):
    """At the default OGFS settings lambda_G = 0.13 and, for every opacity <= 1,
    the energy radius sqrt(2 ln(o/lambda_G)) sigma < 2.1 sigma, so
    r = 0.7 * 3 sigma = 2.1 sigma and radius_px = ceil(2.1 sigma + 1)."""
    lam = math.exp(orc.ogfs_log_lambda(0.2, 0.2))
    raw = 1 - (1 - 1 / 255) ** 256
    assert abs(lam - (1 / 255 + 0.2 * (raw - 1 / 255))) < 1e-6
    assert math.sqrt(2 * math.log(1 / lam)) < 2.1
    rng = np.random.default_rng(5)
    n = 400
    cam = make_camera(320, 240, center=(0.0, -6.0, 0.0), fovx_deg=60.0)
    means = _front_points(n, rng)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    scales = np.exp(rng.normal(-2.5, 0.6, size=(n, 3)))
    opac = rng.uniform(0.01, 1.0, n)
    out = orc.preprocess(orc_settings(orc, cam, variant=0), means, opac, scales, q,
                         colors=np.full((n, 3), 0.5), mode="f64")
    ok = out["radii"] > 0
    sig = _sigma_px(out["conic_opacity"][ok])
    expect = np.maximum(1, np.ceil(2.1 * sig + 1))
    close = np.abs(2.1 * sig + 1 - np.round(2.1 * sig + 1)) < 1e-6
    assert np.array_equal(out["radii"][ok][~close], expect[~close].astype(np.int32))


def test_ogfs_radius_render_yml_values(orc):
    """render.yml: lambda_G 0.2 (apps/render.py:24-37 conversion), floor 0.01,
    ratio_min 0.1: r = sqrt(2 ln 5) sigma = 1.794 sigma at o = 1 and 0.3 sigma
    for o <= 0.2 (energy floored at 0.01, then the 0.1*3 sigma floor)."""
    eps = 1 / 255
    raw = 1 - (1 - eps) ** 256
    eps_scale = (0.2 - eps) / (raw - eps)
    cam = make_camera(400, 400, center=(0.0, -6.0, 0.0), fovx_deg=60.0)
    n = 50
    rng = np.random.default_rng(6)
    means = _front_points(n, rng) * 0.5
    scales = np.full((n, 3), 0.2)
    q = np.tile([1.0, 0, 0, 0], (n, 1))
    for o, factor in ((1.0, math.sqrt(2 * math.log(5.0))), (0.15, 0.3)):
        s = orc_settings(orc, cam, variant=0, ogfs_epsilon_scale=eps_scale, ogfs_epsilon_max=0.2,
                         ogfs_energy_floor=0.01, ogfs_radius_ratio_min=0.1)
        out = orc.preprocess(s, means, np.full(n, o), scales, q, colors=np.full((n, 3), 0.5),
                             mode="f64")
        sig = _sigma_px(out["conic_opacity"])
        expect = np.maximum(1, np.ceil(factor * sig + 1)).astype(np.int32)
        assert np.array_equal(out["radii"], expect), (o, out["radii"][:5], expect[:5])


def _cube_cameras():
    cams = []
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for d in dirs:
        up = (0, 0, 1) if abs(d[2]) < 0.5 else (0, 1, 0)
        cams.append((np.array(d, float), make_camera(256, 256, center=(0, 0, 0), target=d,
                                                     fovx_deg=100.0, fovy_deg=100.0, up=up)))
    return cams


def _sh_basis_at(orc, dirs, deg=3):
    """Y_k(dir) for k < (deg+1)^2 read out of the oracle's colour: unit
    coefficient 0.1 on basis k for channel 0, so rgb0 = 0.1*Y_k + 0.5 > 0."""
    n = dirs.shape[0]
    M = (deg + 1) ** 2
    Y = np.zeros((n, M))
    axes = np.array([c[0] for c in _cube_cameras()])
    face = np.argmax(dirs @ axes.T, axis=1)
    for f, (_, cam) in enumerate(_cube_cameras()):
        sel = np.where(face == f)[0]
        if sel.size == 0:
            continue
        pts = dirs[sel] * 3.0
        s = orc_settings(orc, cam, sh_degree=deg, variant=1)
        for k in range(M):
            sh = np.zeros((sel.size, M, 3))
            sh[:, k, 0] = 0.1
            out = orc.preprocess(s, pts, np.full(sel.size, 0.5), np.full((sel.size, 3), 0.01),
                                 np.tile([1.0, 0, 0, 0], (sel.size, 1)), shs=sh, mode="f64")
            assert (out["radii"] > 0).all()
            Y[sel, k] = (out["rgb"][:, 0] - 0.5) / 0.1
    return Y


def test_sh_addition_theorem(orc):
    """Sum_m Y_lm(d)^2 = (2l+1)/(4 pi) at every direction (Unsold's theorem)."""
    rng = np.random.default_rng(7)
    d = rng.normal(size=(600, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    Y = _sh_basis_at(orc, d)
    for l in range(4):
        blk = Y[:, l * l:(l + 1) * (l + 1)]
        # the constants are float literals (auxiliary.h:22-39): ~1e-7 relative
        np.testing.assert_allclose((blk ** 2).sum(1), (2 * l + 1) / (4 * math.pi), rtol=1e-6)


def test_sh_orthonormal_quadrature(orc):
    """Gram matrix of the 16 basis functions over the sphere is the identity
    (Fibonacci-lattice quadrature)."""
    n = 6000
    k = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * k / n)
    th = math.pi * (1 + 5 ** 0.5) * k
    d = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    Y = _sh_basis_at(orc, d)
    G = 4 * math.pi / n * (Y.T @ Y)
    np.testing.assert_allclose(G, np.eye(16), atol=5e-3)


def test_sh_degree0_and_clamp(orc):
    cam = make_camera(64, 64, center=(0, -4, 0), fovx_deg=60.0)
    sh = np.zeros((3, 16, 3))
    sh[:, 0, :] = [[1.0, -3.0, 0.0], [0.2, 0.1, -0.5], [-2.0, 0.0, 2.0]]
    s = orc_settings(orc, cam, sh_degree=0, variant=1)
    out = orc.preprocess(s, np.zeros((3, 3)), np.full(3, 0.5), np.full((3, 3), 0.1),
                         np.tile([1.0, 0, 0, 0], (3, 1)), shs=sh, mode="f64")
    v = 0.28209479177387814 * sh[:, 0, :].astype(np.float32).astype(np.float64) + 0.5
    np.testing.assert_allclose(out["rgb"], np.maximum(v, 0), rtol=1e-7)
    assert np.array_equal(out["clamped"].astype(bool), v < 0)


def test_near_plane_cull_and_f32_f64_structure(orc):
    """z_view <= 0.2 is culled; the f32 emulation and f64 agree on the radius
    of well-conditioned Gaussians."""
    cam = make_camera(320, 240, center=(0.0, -6.0, 0.0), fovx_deg=60.0)
    rng = np.random.default_rng(8)
    n = 500
    means = rng.uniform(-3, 3, size=(n, 3))
    means[:50, 1] = -6.0 + rng.uniform(-0.5, 0.19, 50)   # at or behind the near limit
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    sc = np.exp(rng.normal(-2.5, 0.5, size=(n, 3)))
    s = orc_settings(orc, cam, variant=0)
    a = orc.preprocess(s, means, np.full(n, 0.6), sc, q, colors=np.full((n, 3), 0.5), mode="f32")
    b = orc.preprocess(s, means, np.full(n, 0.6), sc, q, colors=np.full((n, 3), 0.5), mode="f64")
    assert (a["radii"][:50] == 0).all()
    same = (a["radii"] == b["radii"]).mean()
    assert same > 0.99
