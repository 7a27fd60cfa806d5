"""Training-side neighbours of the LoG render path on the gfx950 kernels.

* ``SparseOptimizer`` -- LoG/model/sparse_optimizer.py:116-195 with the same
  constructor, buffers and ``step(model, index, params, flag_vis)``: per
  parameter one kernel (``fgs_sparse_adam``) reads the visible rows' Adam state
  in place, updates it and writes the parameter rows, instead of gathering the
  state into copies, ten element-wise passes and scattering it back.  The
  state lives on the device (the reference can keep it on the host; with
  288 GB of HBM it stays next to the parameters).
* ``Counter`` -- LoG/model/counter.py:4-83 with ``update_by_output`` as two
  kernels per render (``fgs_counter_update``).
* ``get_expon_lr_func`` -- the learning-rate schedule (sparse_optimizer.py:6-39).
* ``render_to_check`` -- the point-weight probe of LoG.prepare
  (level_of_gaussian.py:205-219) as a forward without statistics or autograd.
* ``activate_root_return`` / ``get_all`` -- LoG's training gather
  (level_of_gaussian.py:260-294) with Activation.activate_root_return
  (LoG/model/activation.py:27-43) as one forward and one backward kernel
  (``fgs_activate_gather`` on the rows in place, ``fgs_activate_backward``)
  instead of the exp / sigmoid / normalize / SH chains of element-wise torch
  kernels and their autograd graph.
"""
from __future__ import annotations

import math

import torch
import torch.nn as nn

from . import _lib
from ._lib import call, ptr


def get_expon_lr_func(lr_init, lr_final, lr_delay_steps=0, lr_delay_mult=1.0, max_steps=1000000):
    """Learning-rate schedule with the signature of sparse_optimizer.py:6-39.

    The rate moves geometrically from lr_init at step 0 to lr_final at
    max_steps (log lr linear in step / max_steps, clipped to [0, 1]).  With
    lr_delay_steps > 0 it is scaled by a warm-up factor that rises from
    lr_delay_mult to 1 along a quarter sine period over the first
    lr_delay_steps steps.  A negative step, or lr_init = lr_final = 0, gives 0.
    """
    log_a = math.log(lr_init) if lr_init > 0 else -math.inf
    log_b = math.log(lr_final) if lr_final > 0 else -math.inf

    def rate(step):
        if step < 0 or (lr_init == 0.0 and lr_final == 0.0):
            return 0.0
        u = min(max(step / max_steps, 0.0), 1.0)
        lr = math.exp(log_a * (1 - u) + log_b * u)
        if lr_delay_steps <= 0:
            return lr
        ramp = math.sin(0.5 * math.pi * min(max(step / lr_delay_steps, 0.0), 1.0))
        return (lr_delay_mult + (1 - lr_delay_mult) * ramp) * lr

    return rate


class BufferDict(nn.Module):
    """A name -> tensor mapping held as module buffers, so a state_dict
    stores each tensor under ``<attribute>.<name>`` (the optimiser-state
    layout of sparse_optimizer.py:96-120).  ``keys`` lists the names in
    insertion order; ``device`` is the device of the first tensor."""

    def __init__(self, tensors):
        super().__init__()
        self.keys = []
        for name, t in tensors.items():
            self.register_buffer(name, t)
            self.keys.append(name)

    @property
    def device(self):
        return self._buffers[self.keys[0]].device

    def __getitem__(self, name):
        return self._buffers[name]

    def __contains__(self, name):
        return name in self.keys

    def __len__(self):
        return len(self.keys)

    def __repr__(self):
        return f'BufferDict: {self.keys}'

    def items(self):
        return ((name, self._buffers[name]) for name in self.keys)

    def get(self, name, default_value=None):
        return self._buffers[name] if name in self.keys else default_value


class SparseOptimizer(nn.Module):
    """Adam over the visible rows only, with the constructor, buffers and
    state_dict layout of sparse_optimizer.py:122-155 (global_steps, then the
    moment dicts exp_avg / exp_avg_sq / max_exp_avg_sq and the per-row step
    counts ``steps``) and its learning-rate schedules for xyz and scaling."""

    def __init__(self, optimize_keys, lr_dict, model, device=None, xyz_scale=None, use_amsgrad=False) -> None:
        super().__init__()
        device = model.xyz.device if device is None else device
        self.device = device
        self.optimize_keys = optimize_keys
        self.lr_dict = lr_dict
        self.use_amsgrad = use_amsgrad
        self.register_buffer('global_steps', torch.zeros((), dtype=torch.float32, device=device))
        shapes = {k: getattr(model, k) for k in optimize_keys}
        self.state_keys = ['exp_avg', 'exp_avg_sq'] + (['max_exp_avg_sq'] if use_amsgrad else [])
        for moment in self.state_keys:
            setattr(self, moment, BufferDict({k: torch.zeros_like(p, device=device) for k, p in shapes.items()}))
        self.steps = BufferDict({k: torch.zeros((p.shape[0],), dtype=torch.int32, device=device)
                                 for k, p in shapes.items()})
        if xyz_scale is not None:
            horizon = int(lr_dict['max_steps'])
            self.xyz_lr = lr_dict['xyz'] * xyz_scale
            xyz_end = lr_dict.get('xyz_final', lr_dict['xyz'] * 0.01) * xyz_scale
            self.xyz_scheduler_args = get_expon_lr_func(lr_init=self.xyz_lr, lr_final=xyz_end,
                                                        max_steps=horizon)
            if 'scaling' in lr_dict:
                self.scaling_scheduler_args = get_expon_lr_func(
                    lr_init=lr_dict['scaling'], lr_final=lr_dict.get('scaling_final', lr_dict['scaling']),
                    max_steps=horizon)

    def _lr(self, key, step):
        if key == 'xyz':
            lr = self.xyz_scheduler_args(step)
            self.xyz_lr = lr
            return lr
        if key == 'scaling':
            return self.scaling_scheduler_args(step)
        return self.lr_dict[key]

    def step(self, model, index, params, flag_vis):
        """Adam step of row index[j] for every j with flag_vis[j], from the
        gathered parameters ``params`` and their gradients (eps 1e-15, betas
        0.9 / 0.999, the global step for the bias correction)."""
        self.global_steps += 1
        step = float(self.global_steps.item())
        dev = params['xyz'].device if 'xyz' in params else next(iter(params.values())).device
        _lib.require_gpu(torch.empty(0, device=dev))
        idx = index.to(device=dev, dtype=torch.int64).contiguous()
        flag = flag_vis.to(device=dev, dtype=torch.uint8).contiguous()
        n = int(idx.numel())
        if int(flag.numel()) != n:
            raise ValueError("index and flag_vis differ in length")
        bad = torch.zeros((1,), dtype=torch.int32, device=dev)
        stream = _lib.stream_handle(dev)
        for key, param in params.items():
            if param.grad is None:
                continue
            lr = self._lr(key, step)
            target = getattr(model, key)
            m, v = self.exp_avg[key], self.exp_avg_sq[key]
            vmax = self.max_exp_avg_sq[key] if self.use_amsgrad else None
            for t in (target, m, v) + ((vmax,) if vmax is not None else ()):
                if t.device != dev or t.dtype != torch.float32 or not t.is_contiguous():
                    raise ValueError(f"{key}: parameter and optimiser state must be contiguous float32 on {dev}")
            D = int(target[0].numel()) if target.shape[0] > 0 else 1
            sel = param.data.to(torch.float32).contiguous()
            grad = param.grad.to(torch.float32).contiguous()
            call("fgs_sparse_adam", n, D, idx.data_ptr(), flag.data_ptr(), int(target.shape[0]), target.data.data_ptr(),
                 sel.data_ptr(), grad.data_ptr(), m.data_ptr(), v.data_ptr(), ptr(vmax), float(lr), 0.9, 0.999,
                 1e-15, step, bad.data_ptr(), stream)
        if int(bad.item()):
            raise IndexError("SparseOptimizer.step: index outside the parameter rows")


# Counter buffers in state_dict order with their dtype and initial value
# (LoG/model/counter.py:5-20).
_COUNTER_BUFFERS = (
    ('weights_max', torch.float32, 0), ('weights_sum', torch.float32, 0), ('grad_sum', torch.float32, 0),
    ('radii_max', torch.int16, 0), ('visible_count', torch.int16, 0),
    ('radii_max_max', torch.int32, 0), ('area_sum', torch.int32, 0),
    ('radius3d_min', torch.float32, 1), ('radius3d_max', torch.float32, 1),
    ('create_steps', torch.int32, 0))
# the buffers reset() clears (counter.py:26-30)
_COUNTER_RESET = ('weights_max', 'weights_sum', 'radii_max', 'radii_max_max', 'area_sum', 'grad_sum',
                  'visible_count')


class Counter(nn.Module):
    """Per-point training statistics with LoG's buffers (counter.py:4-83)."""

    def __init__(self, num_points, add_depth=False) -> None:
        super().__init__()
        for name, dtype, init in _COUNTER_BUFFERS:
            self.register_buffer(name, torch.full((num_points,), init, dtype=dtype))

    def get_gradmean(self):
        return self.grad_sum / torch.clamp(self.area_sum, min=1)

    def reset(self, num_points):
        for name in _COUNTER_RESET:
            t = getattr(self, name)
            t.set_(torch.zeros((num_points,), device=t.device, dtype=t.dtype))

    def reset_create_steps(self):
        self.create_steps.fill_(0)

    def update_by_output(self, output, fix_parent=False):
        """counter.py:36-83 for every render of ``output``; a render with no
        visible rows, no point ids or a point id past its rows is skipped, as
        there."""
        N = int(self.area_sum.shape[0])
        for i in range(len(output['render'])):
            vf = output['visibility_flag'][i]
            visible_index = vf['index']
            if 'index_node' in vf:
                visible_index = torch.cat([visible_index, vf['index_node']])
            if visible_index.numel() == 0:
                continue
            radii = output['radii'][i]
            flag_vis = radii > 0
            vf['flag_vis'] = flag_vis
            vf['index_vis'] = torch.where(flag_vis)[0]
            point_id = output['point_id'][i]
            if point_id.numel() == 0:
                continue
            if int(point_id.max()) >= visible_index.numel():
                continue
            dev = self.area_sum.device
            _lib.require_gpu(self.area_sum)
            vi = visible_index.to(device=dev, dtype=torch.int64).contiguous()
            rd = radii.to(device=dev, dtype=torch.int32).contiguous()
            pw = output['point_weight'][i].data.to(device=dev, dtype=torch.float32).contiguous()
            pid = point_id.to(device=dev, dtype=torch.int64).contiguous()
            pc = output['point_count'][i].to(device=dev, dtype=torch.int64).contiguous()
            g = output['viewspace_points'][i].grad
            gr = g.to(device=dev, dtype=torch.float32).contiguous() if g is not None else None
            if gr is not None and (gr.dim() != 2 or gr.shape[1] != 3):
                raise ValueError("viewspace gradient must be [n, 3]")
            bad = torch.zeros((1,), dtype=torch.int32, device=dev)
            call("fgs_counter_update", int(vi.numel()), vi.data_ptr(), rd.data_ptr(), pw.data_ptr(),
                 int(pid.numel()), pid.data_ptr(), pc.data_ptr(), ptr(gr), N, self.create_steps.data_ptr(),
                 self.visible_count.data_ptr(), self.weights_max.data_ptr(), self.weights_sum.data_ptr(),
                 self.radii_max.data_ptr(), self.radii_max_max.data_ptr(), self.area_sum.data_ptr(),
                 self.grad_sum.data_ptr(), bad.data_ptr(), _lib.stream_handle(dev))
            if int(bad.item()):
                raise IndexError("Counter.update_by_output: index outside the counters")


@torch.no_grad()
def render_to_check(rasterizer, xyz, scaling, rotation, opacity):
    """point_weight of a white render of the given Gaussians
    (LoG/model/level_of_gaussian.py:205-219).  The reference goes through the
    rasterizer module, which also collects RenderStats (on by default); only
    the per-Gaussian maximum weight is used, so the forward here runs without
    statistics and without an autograd record."""
    from . import ext
    rs = rasterizer.raster_settings
    P = int(xyz.shape[0])
    if P == 0:
        return torch.zeros((0,), dtype=torch.float32, device=xyz.device)
    out = ext.rasterize_gaussians(
        rs.bg, xyz, torch.ones_like(xyz), opacity, scaling, rotation, rs.scale_modifier, torch.Tensor([]),
        rs.viewmatrix, rs.projmatrix, rs.tanfovx, rs.tanfovy, rs.image_height, rs.image_width,
        torch.Tensor([]), rs.sh_degree, rs.campos, rs.prefiltered, True, rs.debug,
        ogfs_epsilon_scale=rs.ogfs_epsilon_scale, ogfs_epsilon_max=rs.ogfs_epsilon_max,
        ogfs_energy_floor=rs.ogfs_energy_floor, ogfs_radius_ratio_min=rs.ogfs_radius_ratio_min,
        ogfs_enable_stats=False)
    return out[4]


class _ActivateRows(torch.autograd.Function):
    """Activation.activate_root_return (LoG/model/activation.py:36-43) of
    rows in place: scaling = exp, rotation = F.normalize, opacity = sigmoid,
    colors = SH2RGB(dc) + eval_sh_wobase(dir, rest) for active degree > 0
    with dir from the detached means (activation.py:27-35)."""

    @staticmethod
    def forward(ctx, scaling, rotation, opacity, colors, shs, xyz, campos, deg):
        N = int(scaling.shape[0])
        dev = scaling.device
        rest_coeffs = int(shs.shape[1]) if deg > 0 else 0
        f32 = dict(dtype=torch.float32, device=dev)
        out_s = torch.empty((N, 3), **f32)
        out_r = torch.empty((N, 4), **f32)
        out_o = torch.empty(opacity.shape, **f32)
        out_c = torch.empty((N, 3), **f32)
        if N:
            bad = torch.zeros((1,), dtype=torch.int32, device=dev)
            src = [xyz, scaling, rotation, opacity, colors]
            call("fgs_activate_gather", N, None, N, *[ptr(t) for t in src],
                 ptr(shs) if deg > 0 else None, rest_coeffs, int(deg), ptr(campos) if deg > 0 else None,
                 None, out_s.data_ptr(), out_r.data_ptr(), out_o.data_ptr(), out_c.data_ptr(),
                 bad.data_ptr(), _lib.stream_handle(dev))
        ctx.deg = int(deg)
        ctx.rest_shape = tuple(shs.shape) if shs is not None else None
        ctx.save_for_backward(scaling, rotation, opacity, xyz, campos)
        return out_s, out_r, out_o, out_c

    @staticmethod
    def backward(ctx, g_s, g_r, g_o, g_c):
        scaling, rotation, opacity, xyz, campos = ctx.saved_tensors
        N = int(scaling.shape[0])
        deg = ctx.deg
        dev = scaling.device
        f32 = dict(dtype=torch.float32, device=dev)
        d_s = torch.empty((N, 3), **f32)
        d_r = torch.empty((N, 4), **f32)
        d_o = torch.empty(opacity.shape, **f32)
        d_dc = torch.empty((N, 3), **f32)
        want_rest = deg > 0 and ctx.rest_shape is not None
        d_rest = torch.empty(ctx.rest_shape, **f32) if want_rest else None
        c = lambda t: t.contiguous() if t is not None else None
        g_s, g_r, g_o, g_c = c(g_s), c(g_r), c(g_o), c(g_c)
        if N:
            call("fgs_activate_backward", N, ptr(xyz) if want_rest else None, ptr(scaling), ptr(rotation),
                 ptr(opacity), int(ctx.rest_shape[1]) if want_rest else 0, deg if want_rest else 0,
                 ptr(campos) if want_rest else None, ptr(g_s), ptr(g_r), ptr(g_o), ptr(g_c),
                 d_s.data_ptr(), d_r.data_ptr(), d_o.data_ptr(), d_dc.data_ptr(),
                 d_rest.data_ptr() if want_rest and d_rest.numel() else None, _lib.stream_handle(dev))
        return d_s, d_r, d_o, d_dc, d_rest, None, None, None


def activate_root_return(ret, camera, active_sh_degree):
    """Activation.activate_root_return(ret, camera, active_sh_degree)
    (LoG/model/activation.py:36-43) on the device: same dict in (xyz,
    scaling, opacity, rotation, colors, and shs when the degree uses them),
    same dict out, gradients to the raw rows through one backward kernel."""
    xyz = ret['xyz']
    _lib.require_gpu(xyz)
    deg = int(active_sh_degree) if camera is not None else 0
    shs = ret.get('shs') if deg > 0 else None
    if deg > 0 and shs is None:
        raise ValueError("activate_root_return: SH degree > 0 needs ret['shs']")
    dev = xyz.device
    c = lambda t: t.contiguous() if t is not None else None
    campos = (camera['camera_center'].to(device=dev, dtype=torch.float32).reshape(3).contiguous()
              if deg > 0 else None)
    scaling, rotation, opacity, colors = _ActivateRows.apply(
        c(ret['scaling']), c(ret['rotation']), c(ret['opacity']), c(ret['colors']), c(shs),
        xyz.detach().contiguous(), campos, deg)
    return {'xyz': xyz, 'scaling': scaling, 'opacity': opacity, 'rotation': rotation, 'colors': colors}


def get_all(log_model, camera, rasterizer=None):
    """LoG.get_all(camera, rasterizer) (LoG/model/level_of_gaussian.py:260-294)
    with the activation fused: the visible rows (and with fix_parent the
    visible parents, untrained) are gathered as the reference does, leaves
    become nn.Parameter in training and land in visibility_flag['params'],
    then `activate_root_return` above."""
    g = log_model.gaussian
    vf = g.visibility_flag
    visible_index = vf['index']
    pick = lambda val, idx: nn.Parameter(val[idx]) if log_model.training else val[idx]
    if log_model.fix_parent:
        ret = {key: pick(val, visible_index) for key, val in g.items()}
        vf['params'] = ret
        if 'index_node' in vf:
            index_node = vf['index_node']
            ret = {key: torch.cat([ret[key], val[index_node]]) for key, val in g.items()}
    else:
        if 'index_node' in vf:
            visible_index = torch.cat([visible_index, vf['index_node']])
        ret = {key: pick(val, visible_index) for key, val in g.items()}
        vf['params'] = ret
    return activate_root_return(ret, camera, g.active_sh_degree)
