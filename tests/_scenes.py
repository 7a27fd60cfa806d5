"""Small seeded scenes shared by the oracle and GPU parity tests."""
import numpy as np

from filtergs_amd import workloads
from filtergs_amd.camera import make_camera


def small_scene(n=2000, W=160, H=120, seed=0, half=2.0, dist=8.0, sh_degree=3):
    g = workloads.gaussians(n, half, sh_degree, seed)
    cam = make_camera(W, H, center=(0.0, -dist, 0.5), fovx_deg=60.0)
    return g, cam


def orc_settings(orc, cam, **kw):
    return orc.settings(cam.height, cam.width, cam.tanfovx, cam.tanfovy, cam.world_view_transform,
                        cam.full_proj_transform, cam.center, **kw)
