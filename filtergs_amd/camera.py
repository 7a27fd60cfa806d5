"""Pinhole camera matrices in the convention the rasterizer consumes.

The rasterizer reads ``viewmatrix`` and ``projmatrix`` as flat float[16]
arrays in column-major order (translation at [12..14]).  Callers of the
reference build them row-vector style and transposed: see
``LoG/dataset/base.py:20-55`` (``prepare_camera``) and
``LoG/utils/camera.py:7-28`` (``getProjectionMatrix2``, principal point and
z_ndc in [0, 1] with w = z_view).  This module builds the same matrices for
synthetic cameras (look-at poses, OpenCV axes: x right, y down, z forward).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


def projection_matrix(fx, fy, cx, cy, W, H, znear, zfar, skew=0.0):
    """P_std (4x4, column-vector convention) of getProjectionMatrix2
    (LoG/utils/camera.py:7-28)."""
    P = np.zeros((4, 4), dtype=np.float64)
    P[0, 0] = 2 * fx / W
    P[0, 1] = 2 * skew / W
    P[0, 2] = -1 + 2 * (cx / W)
    P[1, 1] = 2 * fy / H
    P[1, 2] = -1 + 2 * (cy / H)
    P[2, 2] = zfar / (zfar - znear)
    P[2, 3] = -(zfar * znear) / (zfar - znear)
    P[3, 2] = 1.0
    return P


def look_at(center, target, up=(0.0, 0.0, 1.0)):
    """World-to-camera (R, T) with OpenCV axes looking from center to target."""
    c = np.asarray(center, dtype=np.float64)
    z = np.asarray(target, dtype=np.float64) - c
    z /= np.linalg.norm(z)
    u = np.asarray(up, dtype=np.float64)
    x = np.cross(z, u)
    if np.linalg.norm(x) < 1e-9:
        x = np.cross(z, np.array([0.0, 1.0, 0.0]))
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z], axis=0)
    T = -R @ c
    return R, T


@dataclass
class Camera:
    """One view: image size, intrinsics and the matrices of prepare_camera."""
    width: int
    height: int
    fx: float
    fy: float
    cx: float
    cy: float
    R: np.ndarray
    T: np.ndarray
    znear: float = 0.01
    zfar: float = 100.0

    @property
    def tanfovx(self) -> float:
        return self.width / (2.0 * self.fx)

    @property
    def tanfovy(self) -> float:
        return self.height / (2.0 * self.fy)

    @property
    def center(self) -> np.ndarray:
        return (-self.R.T @ self.T).astype(np.float32)

    @property
    def world_view_transform(self) -> np.ndarray:
        """V_std^T (row-vector form), float32; flat memory = column-major V."""
        V = np.eye(4)
        V[:3, :3] = self.R
        V[:3, 3] = self.T
        return V.T.astype(np.float32)

    @property
    def projection(self) -> np.ndarray:
        return projection_matrix(self.fx, self.fy, self.cx, self.cy, self.width, self.height,
                                 self.znear, self.zfar).astype(np.float32).T

    @property
    def full_proj_transform(self) -> np.ndarray:
        """(P_std V_std)^T, computed in float64 then rounded like prepare_camera."""
        V = np.eye(4)
        V[:3, :3] = self.R
        V[:3, 3] = self.T
        return (V.T @ self.projection.astype(np.float64)).astype(np.float32)

    def as_dict(self) -> dict:
        """The camera dict that FilterGS / LoG callers pass around."""
        return dict(image_width=self.width, image_height=self.height,
                    FoVx=2 * math.atan(self.tanfovx), FoVy=2 * math.atan(self.tanfovy),
                    world_view_transform=self.world_view_transform,
                    full_proj_transform=self.full_proj_transform,
                    camera_center=self.center)


def make_camera(width, height, center, target=(0.0, 0.0, 0.0), fovx_deg=60.0, fovy_deg=None,
                up=(0.0, 0.0, 1.0), znear=0.01, zfar=100.0) -> Camera:
    fx = width / (2.0 * math.tan(math.radians(fovx_deg) / 2))
    fy = fx if fovy_deg is None else height / (2.0 * math.tan(math.radians(fovy_deg) / 2))
    R, T = look_at(center, target, up)
    return Camera(width, height, fx, fy, width / 2.0, height / 2.0, R, T, znear, zfar)
