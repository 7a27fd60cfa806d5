"""Seeded synthetic scenes and cameras for the benchmark configurations.

Distributions (shared by every configuration): means uniform in a cube,
log-scales N(-3, 0.5) per axis, unit quaternions from normalised N(0,1)^4,
opacity sigmoid(N(0,1)), SH degree 3 with DC N(0, 0.5) and the other 15
coefficients N(0, 0.05).  Scales are returned activated (exp) because the
rasterizer consumes activated parameters.
"""
from __future__ import annotations

import math

import numpy as np

from .camera import Camera, make_camera


def gaussians(n: int, half_extent: float, sh_degree: int = 3, seed: int = 0,
              dtype=np.float32) -> dict:
    rng = np.random.default_rng(seed)
    M = (sh_degree + 1) ** 2
    means = rng.uniform(-half_extent, half_extent, size=(n, 3))
    log_scales = rng.normal(-3.0, 0.5, size=(n, 3))
    q = rng.normal(0.0, 1.0, size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    opacity = 1.0 / (1.0 + np.exp(-rng.normal(0.0, 1.0, size=(n, 1))))
    shs = np.empty((n, M, 3))
    shs[:, 0, :] = rng.normal(0.0, 0.5, size=(n, 3))
    if M > 1:
        shs[:, 1:, :] = rng.normal(0.0, 0.05, size=(n, M - 1, 3))
    return dict(means3D=means.astype(dtype), log_scales=log_scales.astype(dtype),
                scales=np.exp(log_scales).astype(dtype), rotations=q.astype(dtype),
                opacities=opacity.astype(dtype), shs=shs.astype(dtype))


def camera_c0() -> Camera:
    """800x800, FoV 60 deg, distance 12 from the origin."""
    return make_camera(800, 800, center=(0.0, -12.0, 0.0), fovx_deg=60.0, fovy_deg=60.0)


def camera_c1(width=1920, height=1080, distance=60.0) -> Camera:
    """1080p view of the 40^3 box, FoVx 60 deg."""
    return make_camera(width, height, center=(0.0, -distance, 0.0), fovx_deg=60.0)


def orbit_cameras(n_views: int, seed: int = 3, radius_range=(8.0, 30.0),
                  resolutions=((960, 540), (1920, 1080), (2560, 1440), (3840, 2160))):
    """Orbit poses around the origin at random radius, uniform resolutions."""
    rng = np.random.default_rng(seed)
    cams = []
    for v in range(n_views):
        W, H = resolutions[int(rng.integers(0, len(resolutions)))]
        r = float(rng.uniform(*radius_range))
        theta = 2 * math.pi * v / n_views
        elev = float(rng.uniform(-0.3, 0.5))
        c = (r * math.cos(theta) * math.cos(elev), r * math.sin(theta) * math.cos(elev),
             r * math.sin(elev))
        cams.append(make_camera(W, H, center=c, fovx_deg=60.0))
    return cams


def config(name: str, seed: int = 0):
    """(gaussians, camera) for the named configuration."""
    if name == "c0":
        return gaussians(100_000, 5.0, 3, seed), camera_c0()
    if name == "c1":
        return gaussians(1_000_000, 20.0, 3, seed), camera_c1()
    if name == "c2b":
        return gaussians(3_000_000, 20.0, 3, seed), camera_c1()
    raise KeyError(name)
