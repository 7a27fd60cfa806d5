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
