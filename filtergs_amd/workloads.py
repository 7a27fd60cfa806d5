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
    if name == "c2a":
        # 4096 Gaussians in [-1, 1]^3 seen from distance 4: most of them on
        # screen, 128x128 (the differentiation check)
        return gaussians(4096, 1.0, 3, seed), make_camera(128, 128, center=(0.0, -4.0, 0.0),
                                                          fovx_deg=60.0, fovy_deg=60.0)
    if name == "c2b":
        return gaussians(3_000_000, 20.0, 3, seed), camera_c1()
    raise KeyError(name)


def terrain_height(x, y, seed: int = 0):
    """Seeded 2.5D height field in metres (torch tensors or numpy arrays):
    a sum of a few rotated sinusoids of 1 km to 20 m wavelength."""
    rng = np.random.default_rng(seed + 1000)
    h = 0.0
    for amp, wl in ((45.0, 1100.0), (18.0, 420.0), (6.0, 130.0), (2.0, 37.0), (0.6, 11.0)):
        th = float(rng.uniform(0, math.pi))
        ph = float(rng.uniform(0, 2 * math.pi))
        k = 2 * math.pi / wl
        u = x * math.cos(th) + y * math.sin(th)
        v = -x * math.sin(th) + y * math.cos(th)
        h = h + amp * (u * k + ph).sin() * (0.7 * v * k).cos() if hasattr(u, "sin") else \
            h + amp * np.sin(u * k + ph) * np.cos(0.7 * v * k)
    return h


def lod_tree(levels: int = 12, roots_side: int = 2, extent: float = 2000.0, sh_degree: int = 1,
             seed: int = 0, device="cpu"):
    """Urban-scale LoD tree: a quadtree over an `extent` x `extent` area with
    roots_side^2 roots and `levels` levels (level k is a grid of
    (roots_side * 2^k)^2 nodes, every internal node has 4 children, child
    scale = parent scale / 2).  Leaves sit on the seeded 2.5D height field
    (scaled with the area: its 2 km form times extent / 2 km) with jitter
    inside their cell; an internal node is the mean of its children.  Nodes are numbered level by level, row-major inside a level.

    Returns torch tensors on `device`: means3D [N,3], log_scales [N,3],
    rotations [N,4] (raw), opacities [N,1], shs [N,M,3], tree_node_index [N]
    (own index for internal nodes, -1 for leaves), ancestor_path [N,levels]
    (ancestor at depth d in column d, -1 beyond the node's depth: the
    apps/ancestor.py format), index_parent [N] (-1 at a root), level [N].
    """
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    dev = torch.device(device)
    sides = [roots_side << k for k in range(levels)]
    counts = [s * s for s in sides]
    offs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    N = int(offs[-1])
    L = levels
    # leaves
    sl = sides[-1]
    cell = extent / sl
    iy, ix = torch.meshgrid(torch.arange(sl, dtype=torch.float32), torch.arange(sl, dtype=torch.float32),
                            indexing="ij")
    jit = (torch.rand((sl, sl, 2), generator=g) - 0.5) * 0.8 * cell
    x = (ix + 0.5) * cell - extent / 2 + jit[..., 0]
    y = (iy + 0.5) * cell - extent / 2 + jit[..., 1]
    k = extent / 2000.0   # the height field scales with the area
    z = terrain_height(x / k, y / k, seed) * k + torch.randn((sl, sl), generator=g) * 0.3 * k
    level_means = [torch.stack([x, y, z], -1)]
    for k in range(levels - 2, -1, -1):
        m = level_means[0]
        s = sides[k]
        level_means.insert(0, m.view(s, 2, s, 2, 3).mean(dim=(1, 3)))
    means = torch.cat([m.reshape(-1, 3) for m in level_means])
    log_scales = torch.cat([torch.full((counts[k], 3), math.log(0.5 * extent / sides[k])) for k in range(levels)])
    log_scales = log_scales + torch.randn((N, 3), generator=g) * 0.15
    q = torch.randn((N, 4), generator=g)
    opac = torch.sigmoid(torch.randn((N, 1), generator=g) + 1.0)
    M = (sh_degree + 1) ** 2
    shs = torch.randn((N, M, 3), generator=g) * 0.05
    shs[:, 0, :] = torch.randn((N, 3), generator=g) * 0.5
    # structure
    level = torch.cat([torch.full((counts[k],), k, dtype=torch.int32) for k in range(levels)])
    idx = torch.arange(N, dtype=torch.int64)
    node_index = torch.where(level < levels - 1, idx.to(torch.int32), torch.full_like(level, -1))
    anc = torch.full((N, L), -1, dtype=torch.int32)
    for k in range(1, levels):
        s = sides[k]
        local = torch.arange(counts[k], dtype=torch.int64)
        ly, lx = local // s, local % s
        for d in range(k):
            sh = k - d
            ad = (ly >> sh) * sides[d] + (lx >> sh) + int(offs[d])
            anc[int(offs[k]):int(offs[k + 1]), d] = ad.to(torch.int32)
    parent = torch.full((N,), -1, dtype=torch.int32)
    lv = level.to(torch.int64)
    nz = lv > 0
    parent[nz] = anc[nz, lv[nz] - 1]
    out = dict(means3D=means, log_scales=log_scales, rotations=q, opacities=opac, shs=shs,
               tree_node_index=node_index, ancestor_path=anc, index_parent=parent, level=level)
    return {k: v.contiguous().to(dev) for k, v in out.items()}


def children_table(index_parent, max_child):
    """TensorTree layout (LoG/model/tensor_tree.py:4-26) of a tree given by
    parent links: node_index [N] int32 (the node's row in `tree`, -1 for a
    node without children) and tree [rows, max_child] int32 (child ids in
    ascending id order, -1 in unused slots).  Rows follow the parents' ids."""
    import torch
    parent = index_parent.to(torch.int64)
    N = parent.numel()
    dev = parent.device
    has_child = torch.zeros(N, dtype=torch.bool, device=dev)
    nz = parent >= 0
    has_child[parent[nz]] = True
    rank = torch.cumsum(has_child.to(torch.int64), 0) - 1
    node_index = torch.where(has_child, rank, torch.full_like(rank, -1)).to(torch.int32)
    child = torch.nonzero(nz).reshape(-1)
    order = torch.argsort(parent[child] * N + child)
    child = child[order]
    par = parent[child]
    first = torch.ones_like(par, dtype=torch.bool)
    first[1:] = par[1:] != par[:-1]
    start = torch.cummax(torch.where(first, torch.arange(par.numel(), device=dev), torch.zeros_like(par)), 0).values
    slot = torch.arange(par.numel(), device=dev) - start
    if slot.numel() and int(slot.max()) >= max_child:
        raise ValueError("a node has more than max_child children")
    tree = torch.full((int(has_child.sum()), max_child), -1, dtype=torch.int32, device=dev)
    tree[rank[par], slot] = child.to(torch.int32)
    return node_index, tree


def aerial_cameras(n_views: int = 8, extent: float = 2000.0, seed: int = 5,
                   altitude_range=(150.0, 400.0), width: int = 1920, height: int = 1080):
    """Oblique aerial views over the LoD terrain: positions uniform over the
    central 60 % of the area at altitude U(altitude_range), each looking at a
    ground point 150-400 m ahead, FoVx 60 deg, far plane 5 km."""
    rng = np.random.default_rng(seed)
    cams = []
    for _ in range(n_views):
        cx, cy = rng.uniform(-0.3 * extent, 0.3 * extent, 2)
        alt = float(rng.uniform(*altitude_range))
        th = float(rng.uniform(0, 2 * math.pi))
        d = float(rng.uniform(150.0, 400.0))
        tx, ty = cx + d * math.cos(th), cy + d * math.sin(th)
        ground = float(terrain_height(np.array(cx), np.array(cy)))
        cams.append(make_camera(width, height, center=(float(cx), float(cy), ground + alt),
                                target=(tx, ty, float(terrain_height(np.array(tx), np.array(ty)))),
                                fovx_deg=60.0, up=(0.0, 0.0, 1.0), znear=0.5, zfar=5000.0))
    return cams


def ssim_window(size: int = 11, sigma: float = 1.5):
    """Normalised 2D Gaussian window (outer product of the 1D one)."""
    import torch
    x = torch.arange(size, dtype=torch.float64) - size // 2
    w = torch.exp(-x * x / (2.0 * sigma * sigma))
    w = w / w.sum()
    return (w[:, None] * w[None, :]).to(torch.float32)


def l1_ssim_loss(render, target, window=None):
    """The training objective of configuration 2b, LoG's photometric loss
    (LoG/render/renderer.py:260-273 with LoG/render/loss.py:6-44):
    0.8 * mean |render - target| + 0.2 * (1 - mean SSIM), SSIM over 11x11
    Gaussian windows (sigma 1.5) on the valid region (no padding), per
    channel, C1 = 0.01^2, C2 = 0.03^2.  render, target: [3, H, W] or
    [B, 3, H, W] float tensors on one device."""
    import torch
    import torch.nn.functional as F
    x = render if render.dim() == 4 else render[None]
    y = target if target.dim() == 4 else target[None]
    C = x.shape[1]
    w = (window if window is not None else ssim_window()).to(device=x.device, dtype=x.dtype)
    k = w.expand(C, 1, *w.shape)

    def blur(t):
        return F.conv2d(t, k, groups=C)

    mx, my = blur(x), blur(y)
    sxx = blur(x * x) - mx * mx
    syy = blur(y * y) - my * my
    sxy = blur(x * y) - mx * my
    c1, c2 = 0.01 ** 2, 0.03 ** 2
    ssim = ((2 * mx * my + c1) * (2 * sxy + c2)) / ((mx * mx + my * my + c1) * (sxx + syy + c2))
    return 0.8 * (x - y).abs().mean() + 0.2 * (1.0 - ssim.mean())


def band_camera(cam, y0: int, rows: int):
    """The camera of rows [y0, y0 + rows) of `cam`'s image (same pose and
    focal lengths, principal point moved up by y0)."""
    return Camera(cam.width, rows, cam.fx, cam.fy, cam.cx, cam.cy - y0, cam.R, cam.T, cam.znear,
                  cam.zfar)
