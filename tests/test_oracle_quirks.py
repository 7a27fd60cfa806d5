"""Pins the reference-mode epsilon of the oracle's cov2D backward
(antialias/cuda_rasterizer/backward.cu:203, `denom2inv = 1 / (denom^2 + 1e-7)`) by a consequence the
exact mode does not share.  In the dilate variant both modes lift the 2D
covariance by 0.3 on the diagonal, and with only dL/dconic non-zero every
output of the preprocess backward (dL/dcov3D, dL/dmean3D, dL/dscale,
dL/drot) is linear in (dL/da, dL/db, dL/dc), which the two modes scale by
1/denom^2 and 1/(denom^2 + 1e-7).  So reference / exact = denom^2 /
(denom^2 + 1e-7) for every output of a Gaussian, where denom = det of the
lifted 2D covariance = 1 / det(conic) of the forward.  An epsilon dropped,
added to denom instead of denom^2, or applied in the exact mode fails this."""
import numpy as np

from tests._scenes import orc_settings, small_scene


def test_reference_mode_cov2d_epsilon_ratio(orc):
    g, cam = small_scene(n=200, W=48, H=40, seed=61, dist=5.0, half=1.2)
    p = {k: g[k].astype(np.float64) for k in ("means3D", "scales", "rotations", "opacities")}
    P = p["means3D"].shape[0]
    # small scales: the lifted determinant sits near 0.09, where the epsilon
    # moves the gradient by about 1e-5
    p["scales"] = p["scales"] * 0.02
    colors = np.full((P, 3), 0.5)
    outs = {}
    for mode in (0, 1):
        s = orc_settings(orc, cam, variant=1, grad_mode=mode)
        pre = orc.preprocess(s, p["means3D"], p["opacities"].reshape(-1), p["scales"], p["rotations"],
                             colors=colors, mode="f64")
        rng = np.random.default_rng(62)
        dconic = rng.normal(size=(P, 3))
        zeros = np.zeros((P, 3))
        outs[mode] = (pre, orc.preprocess_bwd(s, p["means3D"], pre["radii"], pre["cov3D"], zeros, dconic,
                                              np.zeros(P), zeros, scales=p["scales"],
                                              rotations=p["rotations"],
                                              opacities=p["opacities"].reshape(-1), mode="f64"))
    pre, ref = outs[0]
    _, exact = outs[1]
    assert np.array_equal(pre["radii"], outs[1][0]["radii"])
    vis = pre["radii"] > 0
    assert vis.sum() > 50
    co = pre["conic_opacity"][vis]
    denom = 1.0 / (co[:, 0] * co[:, 2] - co[:, 1] ** 2)
    assert denom.min() > 0.089 and denom.min() < 0.2
    want = denom ** 2 / (denom ** 2 + 1e-7)
    assert (1.0 - want).max() > 5e-6          # the epsilon is visible at this size
    checked = 0
    for k in ("dL_dcov3D", "dL_dmeans3D", "dL_dscales", "dL_drot"):
        r, e = ref[k][vis], exact[k][vis]
        big = np.abs(e) > 1e-6 * np.abs(e).max()
        ratio = np.where(big, r / np.where(big, e, 1.0), np.nan)
        w = np.broadcast_to(want[:, None], ratio.shape)
        np.testing.assert_allclose(ratio[big], w[big], rtol=1e-9, err_msg=k)
        checked += int(big.sum())
    assert checked > 1000


def _two_on_axis(orc, variant, grad_mode, o_front, o_back=None):
    """One or two Gaussians on the optical axis of a 33x33 view, so that they
    project exactly onto the centre pixel (16, 16), where G = 1; dL/dpix is
    non-zero there only."""
    from filtergs_amd.camera import make_camera
    cam = make_camera(33, 33, center=(0.0, -5.0, 0.0), target=(0.0, 0.0, 0.0), fovx_deg=40.0)
    means = [[0.0, 0.0, 0.0]] + ([[0.0, 1.0, 0.0]] if o_back is not None else [])
    P = len(means)
    p = dict(means3D=np.array(means, np.float64), scales=np.full((P, 3), 0.05),
             rotations=np.tile([1.0, 0.0, 0.0, 0.0], (P, 1)),
             opacities=np.array([o_front] + ([o_back] if o_back is not None else []), np.float64))
    colors = np.array([[0.9, 0.2, 0.6], [0.1, 0.7, 0.3]][:P], np.float64)
    bg = (0.25, 0.5, 0.125)   # exact in float32 (the settings hold floats)
    s = orc_settings(orc, cam, variant=variant, grad_mode=grad_mode, bg=bg)
    f = orc.render(s, p["means3D"], p["opacities"], p["scales"], p["rotations"], colors=colors,
                   mode="f64")
    assert np.array_equal(f["pre32"]["means2D"], np.full((P, 2), 16.0, np.float32))
    dL = np.zeros((3, 33, 33))
    gvec = np.array([0.7, -1.3, 0.4])
    dL[:, 16, 16] = gvec
    ob = orc.backward(s, f, dL, p["means3D"], p["opacities"], p["scales"], p["rotations"],
                      colors=colors, mode="f64")
    return f, ob, colors, np.array(bg), gvec


def test_alpha_clamp_closed_form(orc):
    """The 0.99 clamp of alpha = min(0.99, o*G) (forward.cu:436, backward.cu:499,538).
    One Gaussian on the centre pixel with o*G = 0.999: the image there is
    0.99*c + 0.01*bg.  The reference backward passes the gradient through the
    clamp: dL/do = G * (c - bg) . dL/dpix.  The exact mode differentiates the
    clamped forward, where alpha does not depend on o: dL/do = 0.  Two
    Gaussians, the front one clamped: the front's reference gradient is the
    unclamped derivative of C = c1 a1 + c2 a2 (1 - a1) + bg (1 - a1)(1 - a2),
    (c1 - c2 a2 - bg (1 - a2)) . g, and 0 in the exact mode; the back one
    gets (1 - a1)(c2 - bg) . g in both."""
    a = float(np.float32(0.99))   # the clamp is the float literal 0.99f in both precisions
    for variant in (0, 1):
        f, ref, c, bg, g = _two_on_axis(orc, variant, 0, 0.999)
        np.testing.assert_allclose(f["color"][:, 16, 16], a * c[0] + (1 - a) * bg, rtol=1e-12)
        np.testing.assert_allclose(ref["dL_dopacity"].reshape(-1)[0], np.dot(c[0] - bg, g), rtol=1e-12)
        np.testing.assert_allclose(ref["dL_dcolors"][0], a * g, rtol=1e-12)
        _, ex, _, _, _ = _two_on_axis(orc, variant, 1, 0.999)
        assert ex["dL_dopacity"].reshape(-1)[0] == 0.0
        np.testing.assert_allclose(ex["dL_dcolors"][0], a * g, rtol=1e-12)
        # below the clamp both modes agree with the closed form G (c - bg) . g
        _, r2, _, _, _ = _two_on_axis(orc, variant, 0, 0.6)
        _, e2, _, _, _ = _two_on_axis(orc, variant, 1, 0.6)
        for o in (r2, e2):
            np.testing.assert_allclose(o["dL_dopacity"].reshape(-1)[0], np.dot(c[0] - bg, g), rtol=1e-12)

        a2 = 0.5
        f, ref, c, bg, g = _two_on_axis(orc, variant, 0, 0.999, a2)
        want_img = a * c[0] + (1 - a) * a2 * c[1] + (1 - a) * (1 - a2) * bg
        np.testing.assert_allclose(f["color"][:, 16, 16], want_img, rtol=1e-12)
        d = ref["dL_dopacity"].reshape(-1)
        np.testing.assert_allclose(d[0], np.dot(c[0] - c[1] * a2 - bg * (1 - a2), g), rtol=1e-10)
        np.testing.assert_allclose(d[1], (1 - a) * np.dot(c[1] - bg, g), rtol=1e-10)
        _, ex, _, _, _ = _two_on_axis(orc, variant, 1, 0.999, a2)
        d = ex["dL_dopacity"].reshape(-1)
        assert d[0] == 0.0
        np.testing.assert_allclose(d[1], (1 - a) * np.dot(c[1] - bg, g), rtol=1e-10)
