"""The import-name shims resolve the names the reference's callers import, with
the reference's signatures (no GPU needed)."""
import importlib
import inspect
import sys

import pytest


@pytest.fixture()
def shims():
    from filtergs_amd import compat
    compat.install()
    yield
    compat.uninstall()


def _params(fn):
    return [p.name for p in inspect.signature(fn).parameters.values()]


def test_wodilate_names_and_signatures(shims):
    from diff_gaussian_rasterization_wodilate import (GaussianRasterizationSettings,
                                                      GaussianRasterizer)
    from filtergs_amd import rasterizer
    assert GaussianRasterizer is rasterizer.GaussianRasterizer
    # diff_gaussian_rasterization_wodilate/__init__.py:193-210
    fields = list(GaussianRasterizationSettings._fields)
    assert fields[:17] == ["image_height", "image_width", "tanfovx", "tanfovy", "bg",
                           "scale_modifier", "viewmatrix", "projmatrix", "sh_degree", "campos",
                           "prefiltered", "debug", "ogfs_epsilon_scale", "ogfs_epsilon_max",
                           "ogfs_energy_floor", "ogfs_radius_ratio_min", "ogfs_enable_stats"]
    d = GaussianRasterizationSettings._field_defaults
    assert (d["ogfs_epsilon_scale"], d["ogfs_epsilon_max"], d["ogfs_energy_floor"],
            d["ogfs_radius_ratio_min"], d["ogfs_enable_stats"]) == (0.2, 0.2, 0.05, 0.7, True)
    # :245-262
    assert _params(GaussianRasterizer.forward)[:15] == [
        "self", "means3D", "means2D", "opacities", "shs", "colors_precomp", "scales", "rotations",
        "cov3D_precomp", "use_filter", "ogfs_epsilon_scale", "ogfs_epsilon_max",
        "ogfs_energy_floor", "ogfs_radius_ratio_min", "ogfs_enable_stats"]
    assert _params(GaussianRasterizer.compute_radius) == ["self", "positions", "scales",
                                                          "rotations"]
    assert _params(GaussianRasterizer.markVisible) == ["self", "positions"]
    from diff_gaussian_rasterization_wodilate import _C
    assert _C.rasterize_gaussians is importlib.import_module("filtergs_amd.ext").rasterize_gaussians


def test_dilate_names(shims):
    from diff_gaussian_rasterization import GaussianRasterizationSettings, GaussianRasterizer
    from filtergs_amd import dilate
    assert GaussianRasterizer is dilate.GaussianRasterizer
    assert "ogfs_epsilon_scale" not in GaussianRasterizationSettings._fields
    assert _params(GaussianRasterizer.forward)[:9] == [
        "self", "means3D", "means2D", "opacities", "shs", "colors_precomp", "scales", "rotations",
        "cov3D_precomp"]


def test_dilate_C_signatures(shims):
    """The dilate wrapper's _C calls (antialias/diff_gaussian_rasterization/
    __init__.py:60-92, :103-143): rasterize_gaussians takes 19 positional
    arguments and returns 8 values, the backward 21; neither carries the
    wodilate's use_filter / OGFS arguments."""
    from diff_gaussian_rasterization import _C
    from filtergs_amd import dilate_ext, ext
    assert _C is dilate_ext
    assert _C.rasterize_gaussians is not ext.rasterize_gaussians

    def positional(fn):
        return [p.name for p in inspect.signature(fn).parameters.values()
                if p.kind == inspect.Parameter.POSITIONAL_OR_KEYWORD]
    fwd = positional(_C.rasterize_gaussians)
    assert len(fwd) == 19 and fwd[-2:] == ["prefiltered", "debug"]
    assert "use_filter" not in fwd and not any(n.startswith("ogfs") for n in fwd)
    bwd = positional(_C.rasterize_gaussians_backward)
    assert len(bwd) == 21 and bwd[-1] == "debug"
    assert positional(_C.mark_visible) == ["means3D", "viewmatrix", "projmatrix"]
    # the wrapper passes exactly the reference's argument tuples: a fake _C
    # records them and returns 8 forward values
    from filtergs_amd import dilate
    import torch
    seen = {}

    def fake_fwd(*args, **kw):
        seen["fwd"] = (len(args), kw)
        z = torch.zeros(1)
        return 0, z, z, z, torch.zeros(1, dtype=torch.int32), z, z, z

    orig = dilate._C.rasterize_gaussians
    dilate._C.rasterize_gaussians = fake_fwd
    try:
        rs = dilate.GaussianRasterizationSettings(4, 4, 1.0, 1.0, torch.zeros(3), 1.0, torch.eye(4),
                                                  torch.eye(4), 0, torch.zeros(3), False, False)
        out = dilate.GaussianRasterizer(rs)(torch.zeros(1, 3), torch.zeros(1, 3), torch.ones(1, 1),
                                            colors_precomp=torch.ones(1, 3), scales=torch.ones(1, 3),
                                            rotations=torch.ones(1, 4))
    finally:
        dilate._C.rasterize_gaussians = orig
    assert seen["fwd"][0] == 19 and len(out) == 4


def test_filtergs_cuda_names(shims):
    from FilterGS.cuda.compute_radius import compute_radius_module, compute_visibility_radius
    from FilterGS.cuda.gather_attributes import gather_attributes
    from FilterGS.cuda.visible_flag import (visible_flag_cuda, visible_flag_cuda_batch,
                                            visible_flag_cuda_batch_direct,
                                            visible_flag_cuda_direct)
    # FilterGS/cuda/compute_radius.py:23-26
    assert _params(compute_visibility_radius) == [
        "means3d", "scales", "rotations", "projmatrix", "viewmatrix", "tree_node_index",
        "ancestor_path", "focal_x", "focal_y", "tan_fovx", "tan_fovy", "radius_max", "padding",
        "radius_filter_threshold"]
    sig = inspect.signature(compute_visibility_radius).parameters
    assert sig["padding"].default == 0.05 and sig["radius_filter_threshold"].default == 100.0
    assert callable(compute_radius_module.compute_radius)
    assert _params(gather_attributes) == ["index", "xyz", "scaling", "rotation", "opacity",
                                          "colors", "shs"]
    for fn in (visible_flag_cuda, visible_flag_cuda_direct, visible_flag_cuda_batch,
               visible_flag_cuda_batch_direct):
        assert _params(fn)[2:] == ["padding", "use_optimized"]


def test_uninstall_removes_aliases(shims):
    from filtergs_amd import compat
    compat.uninstall()
    assert "diff_gaussian_rasterization_wodilate" not in sys.modules
    assert "FilterGS.cuda.compute_radius" not in sys.modules
    compat.install()


def test_filtergs_class_api(shims):
    """FilterGS/filtergs.py:9-359: constructor, stages and helpers."""
    from filtergs_amd.filtergs import FilterGS
    pkg = sys.modules.get("FilterGS")
    if pkg is not None and getattr(pkg, "__file__", None) is None:   # the shim's package
        from FilterGS import FilterGS as F2
        from FilterGS.filtergs import FilterGS as F3
        assert F2 is FilterGS and F3 is FilterGS
    assert _params(FilterGS.__init__) == ["self", "sh_degree", "xyz_scale"]
    assert _params(FilterGS.prepare) == ["self", "rasterizer", "camera", "tree",
                                         "tree_node_index_gpu", "ancestor_path_gpu", "radius_max",
                                         "padding", "radius_filter_threshold"]
    sig = inspect.signature(FilterGS.prepare).parameters
    assert (sig["radius_max"].default, sig["padding"].default,
            sig["radius_filter_threshold"].default) == (2.0, 0.5, 700.0)
    assert _params(FilterGS.get_all) == ["self", "camera"]
    assert _params(FilterGS.render) == ["self", "camera", "rasterizer", "background", "enable_stats"]
    assert _params(FilterGS.register_by_pointcloud)[:5] == ["self", "xyz", "colors", "scales",
                                                            "init_opacity"]


def test_filtergs_register_by_pointcloud_buffers():
    """register_by_pointcloud stores the inverse activations (filtergs.py:304-348)."""
    import torch
    from filtergs_amd.filtergs import FilterGS
    fg = FilterGS(sh_degree=2)
    xyz = torch.zeros((4, 3))
    colors = torch.full((4, 3), 0.75)
    scales = torch.tensor([1.0, 1.0, 1.0, 1.0])
    fg.register_by_pointcloud(xyz, colors, scales, 0.1)
    assert torch.allclose(fg.scaling, torch.zeros((4, 3)))
    assert torch.allclose(fg.colors, torch.full((4, 3), 0.25 / 0.28209479177387814))
    assert torch.allclose(torch.sigmoid(fg.opacity), torch.full((4, 1), 0.1))
    assert fg.shs.shape == (4, 8, 3) and fg.rotation[:, 0].eq(1).all()


def test_filtergs_register_with_ground_grid():
    """The ground grid of register_by_pointcloud (filtergs.py:322-339,
    :350-362): a grid of spacing init_step over the cloud's xy box widened by
    0.05, at z = height, scales (step, step, step / 10) stored as logs, and
    -- unlike the cloud's points -- colours 0.5 and opacities ground_opacity
    stored without the inverse activations; the per-point scales are clamped
    to [mean / 4, 4 mean] first."""
    import torch
    from filtergs_amd.filtergs import FilterGS
    fg = FilterGS(sh_degree=1)
    xyz = torch.tensor([[0.0, 0.0, 1.0], [1.0, 0.5, 2.0], [0.5, 1.0, 0.0]])
    scales = torch.tensor([0.01, 1.0, 10.0])       # mean 3.67: clamped to [0.9175, 14.67]
    fg.register_by_pointcloud(xyz, torch.full((3, 3), 0.5), scales, 0.5, height=-1.0, init_step=0.5,
                              ground_opacity=0.8)
    m = float(scales.mean())
    s = torch.tensor([m / 4, 1.0, 10.0])
    ax = torch.arange(-0.05, 1.05, 0.5)             # 3 steps per axis over [-0.05, 1.05)
    n = ax.numel() ** 2
    assert fg.xyz.shape == (3 + n, 3)
    assert torch.allclose(fg.scaling[:3], s.log()[:, None].expand(3, 3))
    g = fg.xyz[3:]
    assert torch.allclose(g[:, 2], torch.full((n,), -1.0))
    assert torch.allclose(g[:, 0].reshape(ax.numel(), -1)[:, 0], ax)      # x outer, y inner ("ij")
    assert torch.allclose(g[:, 1].reshape(ax.numel(), -1)[0], ax)
    assert torch.allclose(fg.scaling[3:], torch.log(torch.tensor([0.5, 0.5, 0.05])).expand(n, 3))
    assert torch.allclose(fg.colors[3:], torch.full((n, 3), 0.5))
    assert torch.allclose(fg.opacity[3:], torch.full((n, 1), 0.8))
    assert torch.allclose(fg.opacity[:3], torch.zeros((3, 1)))           # logit(0.5)
    assert fg.shs.shape == (3 + n, 3, 3) and not fg.shs.any()
    assert (fg.rotation == torch.tensor([1.0, 0.0, 0.0, 0.0])).all()
    assert fg.keys[-6:] == ['scaling', 'colors', 'xyz', 'opacity', 'rotation', 'shs']
