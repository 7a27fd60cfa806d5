"""Pins of the oracle's FilterGS activation (LoG/model/activation.py:27-43,
LoG/model/sh_utils.py:31-68): the SH term against the rasterizer oracle's
independently pinned SH colour, the degree-4 basis by quadrature, and the
special values of the element-wise activations."""
import math

import numpy as np

from filtergs_amd.camera import make_camera
from tests._scenes import orc_settings


def test_wobase_colour_equals_rasterizer_sh(orc):
    """SH2RGB(dc) + eval_sh_wobase(dir, rest) equals the rasterizer's SH colour
    computeColorFromSH (forward.cu:48-99) of the same 16 coefficients before
    its clamp, for degrees 1-3."""
    rng = np.random.default_rng(11)
    n = 200
    cam = make_camera(256, 256, center=(0.0, -6.0, 0.0), fovx_deg=60.0)
    pts = np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)], 1)
    sh = np.zeros((n, 16, 3))
    sh[:, 0, :] = rng.uniform(1.0, 2.0, (n, 3))          # keeps the colour positive
    sh[:, 1:, :] = rng.normal(0, 0.05, (n, 15, 3))
    for deg in (1, 2, 3):
        s = orc_settings(orc, cam, sh_degree=deg, variant=1)
        pre = orc.preprocess(s, pts, np.full(n, 0.5), np.full((n, 3), 0.01),
                             np.tile([1.0, 0, 0, 0], (n, 1)), shs=sh, mode="f64")
        assert (pre["radii"] > 0).all()
        act = orc.activate_gather(np.arange(n), pts, np.zeros((n, 3)), np.tile([1.0, 0, 0, 0], (n, 1)),
                                  np.zeros(n), sh[:, 0, :], sh[:, 1:, :], deg, cam.center)
        np.testing.assert_allclose(act["colors"], pre["rgb"], rtol=2e-6)


def test_wobase_degree4_orthonormal():
    """The 24 basis functions of degrees 1-4 are orthonormal on the sphere
    (Fibonacci-lattice quadrature)."""
    from oracle import pyoracle
    n = 8000
    k = np.arange(n) + 0.5
    phi = np.arccos(1 - 2 * k / n)
    th = math.pi * (1 + 5 ** 0.5) * k
    d = np.stack([np.cos(th) * np.sin(phi), np.sin(th) * np.sin(phi), np.cos(phi)], 1)
    B = pyoracle.sh_wobase_basis(d, 4)
    assert B.shape[1] == 24
    G = 4 * math.pi / n * (B.T @ B)
    np.testing.assert_allclose(G, np.eye(24), atol=5e-3)


def test_elementwise_activations(orc):
    idx = np.array([2, 0, 1])
    xyz = np.array([[0.0, 0, 0], [1, 2, 3], [4, 5, 6]])
    logs = np.array([[0.0, 0, 0], [math.log(2.0)] * 3, [-1.0, 0.5, 2.0]])
    rot = np.array([[2.0, 0, 0, 0], [0, 0, 0, 0], [1.0, 1, 1, 1]])
    logit = np.array([0.0, 50.0, -50.0])
    dc = np.zeros((3, 3))
    out = orc.activate_gather(idx, xyz, logs, rot, logit, dc, None, 0, np.zeros(3))
    np.testing.assert_array_equal(out["xyz"], xyz[idx])
    np.testing.assert_allclose(out["scaling"][1], [1.0, 1.0, 1.0])
    np.testing.assert_allclose(out["scaling"][2], [2.0, 2.0, 2.0])
    np.testing.assert_allclose(out["opacity"], [0.0, 0.5, 1.0], atol=1e-20)
    np.testing.assert_allclose(out["rotation"][1], [1.0, 0, 0, 0])
    np.testing.assert_array_equal(out["rotation"][2], [0.0, 0, 0, 0])   # F.normalize of zero
    np.testing.assert_allclose(out["rotation"][0], [0.5] * 4)
    np.testing.assert_allclose(out["colors"], 0.5)                     # SH2RGB(0)
