"""Pins the oracle's backward (blend bwd + preprocess bwd) with FP64 central
finite differences of its f64 forward, in grad_mode 'exact', for both
variants and for the SH / colors_precomp / cov3D_precomp input paths.  The
tile structure is held fixed (it is piecewise constant), and parameters whose
perturbation flips a blend decision (decision hash) are skipped."""
import numpy as np
import pytest

from tests._scenes import orc_settings, small_scene

EPS = 1e-6


class _Problem:
    def __init__(self, orc, variant, path, seed, n=300, W=48, H=40, opac_cap=None, scene=None):
        self.orc = orc
        g, cam = scene if scene is not None else small_scene(n=n, W=W, H=H, seed=seed, dist=5.0,
                                                             half=1.2)
        n, W, H = g["means3D"].shape[0], cam.width, cam.height
        self.s = orc_settings(orc, cam, sh_degree=3, variant=variant, grad_mode=1,
                              bg=(0.2, 0.5, 0.1))
        p = {k: g[k].astype(np.float64) for k in ("means3D", "scales", "rotations", "opacities", "shs")}
        p["opacities"] = p["opacities"].reshape(-1)
        if opac_cap is not None:
            p["opacities"] = np.minimum(p["opacities"], opac_cap)
        rng = np.random.default_rng(seed + 100)
        self.path = path
        if path == "colors":
            p["colors"] = rng.uniform(0.1, 0.9, size=(n, 3))
            del p["shs"]
        if path == "cov3D":
            pre = orc.preprocess(self.s, p["means3D"], p["opacities"], p["scales"], p["rotations"],
                                 colors=np.full((n, 3), 0.5), mode="f64")
            p["cov3D"] = pre["cov3D"].copy()
            del p["scales"], p["rotations"]
        self.p = p
        f32 = orc.preprocess(self.s, **self._args(p, np.float32))
        self.radii = f32["radii"]
        self.keys, self.vals, self.ranges = orc.bin_gaussians(f32["radii"], f32["rect"],
                                                              f32["depths"], W, H)
        self.dL = rng.normal(size=(3, H, W))
        self.base = self.forward(p)

    def _args(self, p, dt):
        a = dict(means3D=p["means3D"].astype(dt), opacities=p["opacities"].astype(dt))
        for k_in, k_out in (("scales", "scales"), ("rotations", "rotations"), ("shs", "shs"),
                            ("colors", "colors"), ("cov3D", "cov3D")):
            if k_in in p:
                a[k_out] = p[k_in].astype(dt)
        a["mode"] = "f32" if dt == np.float32 else "f64"
        return a

    def forward(self, p):
        nt = 1 if p["means3D"].shape[0] < 1000 else 8
        pre = self.orc.preprocess(self.s, **self._args(p, np.float64), nthreads=nt)
        P = pre["radii"].shape[0]
        f = self.orc.blend_fwd(self.s, P, self.ranges, self.vals, pre["means2D"], pre["rgb"],
                               pre["conic_opacity"], mode="f64", decision_hash=True, nthreads=nt)
        f["pre"] = pre
        f["loss"] = float((f["color"] * self.dL).sum())
        return f

    def analytic(self):
        f = self.base
        pre = f["pre"]
        P = pre["radii"].shape[0]
        gb = self.orc.blend_bwd(self.s, P, self.ranges, self.vals, pre["means2D"], pre["rgb"],
                                pre["conic_opacity"], f["final_T"], f["n_contrib"], self.dL,
                                mode="f64")
        p = self.p
        gp = self.orc.preprocess_bwd(self.s, p["means3D"], self.radii, pre["cov3D"],
                                     gb["dL_dmean2D"], gb["dL_dconic"], gb["dL_dopacity"],
                                     gb["dL_dcolors"], p.get("scales"), p.get("rotations"),
                                     p["opacities"], p.get("shs"), mode="f64")
        out = {"means3D": gp["dL_dmeans3D"], "opacities": gp["dL_dopacity"]}
        if "scales" in p:
            out["scales"] = gp["dL_dscales"]
            out["rotations"] = gp["dL_drot"]
        if "shs" in p:
            out["shs"] = gp["dL_dsh"]
        if "colors" in p:
            out["colors"] = gb["dL_dcolors"]
        if "cov3D" in p:
            out["cov3D"] = gp["dL_dcov3D"]
        return out

    def fd(self, name, idx):
        vals = []
        for sgn in (1, -1):
            q = {k: v.copy() for k, v in self.p.items()}
            q[name].reshape(-1)[idx] += sgn * EPS
            f = self.forward(q)
            if not np.array_equal(f["decision_hash"], self.base["decision_hash"]):
                return None
            vals.append(f["loss"])
        return (vals[0] - vals[1]) / (2 * EPS)


def _check(prob, n_per_param=40, seed=0):
    rng = np.random.default_rng(seed)
    ana = prob.analytic()
    vis = np.where(prob.radii > 0)[0]
    checked = 0
    for name, g in ana.items():
        per = g.reshape(g.shape[0], -1).shape[1] if g.ndim > 1 else 1
        cand = rng.choice(vis, size=min(n_per_param, vis.size), replace=False)
        for gi in cand:
            for comp in range(per):
                idx = gi * per + comp
                fd = prob.fd(name, idx)
                if fd is None:
                    continue
                a = g.reshape(-1)[idx]
                scale = max(abs(fd), abs(a), 1e-3)
                assert abs(a - fd) <= 2e-5 * scale + 2e-7, (name, gi, comp, a, fd)
                checked += 1
    return checked


@pytest.mark.parametrize("variant", [0, 1])
def test_fd_exact_sh_path(orc, variant):
    prob = _Problem(orc, variant, "shs", seed=21 + variant)
    assert _check(prob, n_per_param=12) > 300


@pytest.mark.parametrize("variant", [0, 1])
def test_fd_exact_colors_path(orc, variant):
    prob = _Problem(orc, variant, "colors", seed=31 + variant)
    assert _check(prob, n_per_param=15) > 100


def test_fd_exact_cov3d_path(orc):
    prob = _Problem(orc, 1, "cov3D", seed=41)
    assert _check(prob, n_per_param=15) > 100


def test_reference_equals_exact_where_they_coincide(orc):
    """Dilate variant (forward and backward both +0.3), alpha never clamped,
    scale_modifier 1: the reference equations are the exact derivative up to
    the +1e-7 in backward.cu:203."""
    prob = _Problem(orc, 1, "shs", seed=51, opac_cap=0.9)
    ana_exact = prob.analytic()
    prob.s.grad_mode = 0
    ana_ref = prob.analytic()
    for k in ana_exact:
        np.testing.assert_allclose(ana_ref[k], ana_exact[k], rtol=1e-5,
                                   atol=1e-6 * np.abs(ana_exact[k]).max())


@pytest.mark.parametrize("variant", [0, 1])
def test_fd_exact_c2a_size(orc, variant):
    """The differentiation check at its configured size: 4096 Gaussians at
    128x128 (FP64 central differences, eps 1e-6, on a sample of the visible
    Gaussians' means, scales, rotations, opacities and SH)."""
    from filtergs_amd import workloads
    prob = _Problem(orc, variant, "shs", seed=61 + variant,
                    scene=workloads.config("c2a", seed=61 + variant))
    assert (prob.radii > 0).sum() > 3000
    assert _check(prob, n_per_param=6, seed=variant) > 150


@pytest.mark.parametrize("mode", ["f64", "f32"])
def test_threaded_backward_equals_sequential(orc, mode):
    """The threaded backward (tiles and Gaussians split over threads, per-thread
    accumulators) gives the sequential sums up to summation order, and the
    same bits on every run with the same thread count."""
    g, cam = small_scene(n=3000, W=200, H=136, seed=71, dist=5.0)
    s = orc_settings(orc, cam, sh_degree=3, grad_mode=1)
    args = (g["means3D"], g["opacities"], g["scales"], g["rotations"], g["shs"])
    f = orc.render(s, *args, mode=mode)
    dL = np.random.default_rng(72).normal(size=(3, cam.height, cam.width))
    seq = orc.backward(s, f, dL, *args, mode=mode)
    thr = orc.backward(s, f, dL, *args, mode=mode, nthreads=5)
    thr2 = orc.backward(s, f, dL, *args, mode=mode, nthreads=5)
    rtol = 1e-11 if mode == "f64" else 2e-4
    for k in seq:
        a, b = np.asarray(seq[k], np.float64), np.asarray(thr[k], np.float64)
        scale = np.abs(a).max() + 1e-30
        assert np.abs(a - b).max() <= rtol * scale, k
        assert np.array_equal(thr[k], thr2[k]), k
    assert np.abs(seq["dL_dmeans3D"]).max() > 0
