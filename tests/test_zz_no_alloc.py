"""fgs_backward_noalloc -- the backward whose dL/dconic workspace is the
library's static device array -- against fgs_backward with a caller
workspace, through the C ABI: two calls in a row on one device, the second
with fewer Gaussians (the array keeps the first call's contents), and the
size limit."""
import ctypes

import numpy as np
import pytest
import torch

from filtergs_amd import _lib, ext
from tests._scenes import small_scene

pytestmark = pytest.mark.gpu

DEV = "cuda"
MAX_P = 4194304  # FGS_NOALLOC_MAX_GAUSSIANS


def _t(a, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


def _backward(entry, g, cam, dL):
    """Forward through ext, then the named backward entry on fresh outputs."""
    P = g["means3D"].shape[0]
    bg = _t([0.1, 0.2, 0.3])
    vm, pm, cp = _t(cam.world_view_transform), _t(cam.full_proj_transform), _t(cam.center)
    means, sc, ro, sh = _t(g["means3D"]), _t(g["scales"]), _t(g["rotations"]), _t(g["shs"])
    opac = _t(g["opacities"].reshape(P, 1))
    out = ext.rasterize_gaussians(bg, means, torch.Tensor([]), opac, sc, ro, 1.0, torch.Tensor([]), vm, pm,
                                  cam.tanfovx, cam.tanfovy, cam.height, cam.width, sh, 3, cp, False, True,
                                  False, ogfs_enable_stats=False)
    K, color, pid, wpix, pw, radii, geom, binning, img, _ = out
    s, keep_s = ext._settings(bg, vm, pm, cp, cam.height, cam.width, cam.tanfovx, cam.tanfovy, 1.0, 3, False,
                              False, 0, True, 0.2, 0.2, 0.05, 0.7, False, 0, means.device)
    gs, keep_g = ext._gaussians(P, means, sh, torch.Tensor([]), opac.reshape(-1), sc, ro, torch.Tensor([]),
                                means.device)
    f = dict(dtype=torch.float32, device=DEV)
    o = dict(m2=torch.zeros((P, 3), **f), col=torch.zeros((P, 3), **f), op=torch.zeros((P, 1), **f),
             m3=torch.empty((P, 3), **f), cov=torch.empty((P, 6), **f), sh=torch.empty((P, 16, 3), **f),
             sc=torch.empty((P, 3), **f), rot=torch.empty((P, 4), **f))
    dL_t = _t(dL)   # held until the call is done: its memory must not be reused meanwhile
    args = [ctypes.byref(s), ctypes.byref(gs), radii.data_ptr(), geom.data_ptr(), int(K),
            binning.data_ptr(), img.data_ptr(), dL_t.data_ptr()] + [o[k].data_ptr() for k in o]
    stream = _lib.stream_handle(means.device)
    if entry == "fgs_backward":
        ws = torch.empty((_lib.load().fgs_backward_ws_bytes(P),), dtype=torch.uint8, device=DEV)
        # a workspace full of garbage, as a fresh allocation may hold
        ws.fill_(0x7F)
        _lib.call(entry, *args, ws.data_ptr(), stream)
    else:
        _lib.call(entry, *args, stream)
    torch.cuda.synchronize()
    del keep_s, keep_g, dL_t
    tiles = ext.geom_views(geom, P)["tiles"]
    return o, int(tiles.max())


def test_noalloc_matches_workspace_backward():
    scenes = [small_scene(n=6000, W=256, H=192, seed=301, dist=6.0),
              small_scene(n=2500, W=160, H=128, seed=302, dist=5.0)]
    for g, cam in scenes:   # the second call has fewer Gaussians
        P = g["means3D"].shape[0]
        dL = np.random.default_rng(P).normal(size=(3, cam.height, cam.width)).astype(np.float32)
        ref, tmax = _backward("fgs_backward", g, cam, dL)
        got, _ = _backward("fgs_backward_noalloc", g, cam, dL)
        # The four waves of a tile add their per-Gaussian totals in LDS with
        # float atomics, in an order that varies from run to run: the same
        # entry point differs from itself by rounding, so the bar is the one
        # of the blend work shapes (relative L2 1e-5 for the blend terms, 1e-4
        # for the chain through the covariance and rotation).
        for k in ref:
            assert torch.isfinite(got[k]).all()
            r, a = ref[k].double(), got[k].double()
            d = float((a - r).norm() / r.norm().clamp_min(1e-30))
            assert d <= (1e-5 if k in ("m2", "col", "op") else 1e-4), (P, k, d)


def test_noalloc_rejects_more_than_its_workspace():
    g, cam = small_scene(n=10, W=32, H=32, seed=303, dist=5.0)
    s, keep = ext._settings(_t([0.0, 0.0, 0.0]), _t(cam.world_view_transform), _t(cam.full_proj_transform),
                            _t(cam.center), 32, 32, cam.tanfovx, cam.tanfovy, 1.0, 0, False, False, 0, True,
                            0.2, 0.2, 0.05, 0.7, False, 0, torch.device(DEV))
    gs = _lib.Gaussians()
    gs.P = MAX_P + 1
    gs.M = 0
    x = _t(np.zeros(16))
    gs.means3D = gs.opacities = gs.colors_precomp = gs.scales = gs.rotations = x.data_ptr()
    with pytest.raises(_lib.FGSError, match="workspace"):
        _lib.call("fgs_backward_noalloc", ctypes.byref(s), ctypes.byref(gs), *([x.data_ptr()] * 3), 1,
                  *([x.data_ptr()] * 12), _lib.stream_handle(torch.device(DEV)))
