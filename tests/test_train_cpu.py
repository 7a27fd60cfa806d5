"""Host-side pieces of the LoG training neighbours, pinned without a GPU:
the learning-rate schedule by its closed forms, and the state_dict layouts
of SparseOptimizer and Counter (LoG/model/sparse_optimizer.py:6-155,
LoG/model/counter.py:5-36), which checkpoints rely on."""
import math

import torch

from filtergs_amd.train import BufferDict, Counter, SparseOptimizer, get_expon_lr_func


def test_schedule_closed_forms():
    a, b, n = 1.6e-4, 1.6e-6, 30000
    f = get_expon_lr_func(a, b, max_steps=n)
    assert f(0) == a and abs(f(n) - b) < 1e-18 and abs(f(2 * n) - b) < 1e-18
    assert abs(f(n // 2) - math.sqrt(a * b)) < 1e-15           # geometric midpoint
    assert abs(f(n // 4) - a ** 0.75 * b ** 0.25) < 1e-15
    assert f(-1) == 0.0 and get_expon_lr_func(0.0, 0.0)(10) == 0.0
    g = get_expon_lr_func(a, b, lr_delay_steps=100, lr_delay_mult=0.01, max_steps=n)
    assert abs(g(0) - 0.01 * a) < 1e-18                         # warm-up starts at mult
    # half-way through the warm-up: mult + (1 - mult) sin(pi/4)
    assert abs(g(50) - (0.01 + 0.99 * math.sin(math.pi / 4)) * f(50)) < 1e-15
    assert g(100) == f(100) and g(5000) == f(5000)              # warm-up over


class _Model(torch.nn.Module):
    def __init__(self, n):
        super().__init__()
        self.xyz = torch.nn.Parameter(torch.zeros(n, 3))
        self.scaling = torch.nn.Parameter(torch.zeros(n, 3))
        self.colors = torch.nn.Parameter(torch.zeros(n, 3))


def test_optimizer_state_layout():
    m = _Model(7)
    opt = SparseOptimizer(["xyz", "scaling"], {"xyz": 1e-3, "xyz_final": 1e-5, "scaling": 5e-3,
                                               "max_steps": 1000}, m, device="cpu", xyz_scale=2.0,
                          use_amsgrad=True)
    keys = list(opt.state_dict().keys())
    assert keys == ["global_steps"] + [f"{s}.{k}" for s in ("exp_avg", "exp_avg_sq", "max_exp_avg_sq", "steps")
                                       for k in ("xyz", "scaling")]
    assert opt.state_keys == ["exp_avg", "exp_avg_sq", "max_exp_avg_sq"]
    assert opt.exp_avg["xyz"].shape == (7, 3) and opt.steps["scaling"].dtype == torch.int32
    assert opt.exp_avg["xyz"].data_ptr() != opt.exp_avg_sq["xyz"].data_ptr()
    assert opt.xyz_lr == 2e-3 and abs(opt.xyz_scheduler_args(1000) - 2e-5) < 1e-18
    assert abs(opt.scaling_scheduler_args(1000) - 5e-3) < 1e-17
    plain = SparseOptimizer(["colors"], {"colors": 1e-3}, m, device="cpu")
    assert plain.state_keys == ["exp_avg", "exp_avg_sq"] and not hasattr(plain, "xyz_scheduler_args")


def test_buffer_dict_and_counter_layout():
    d = BufferDict({"a": torch.zeros(2), "b": torch.ones(3)})
    assert d.keys == ["a", "b"] and d["b"].sum() == 3 and d.get("c", 5) == 5 and "a" in d
    assert [k for k, _ in d.items()] == ["a", "b"] and list(d.state_dict()) == ["a", "b"]
    c = Counter(4)
    sd = c.state_dict()
    assert list(sd) == ["weights_max", "weights_sum", "grad_sum", "radii_max", "visible_count",
                        "radii_max_max", "area_sum", "radius3d_min", "radius3d_max", "create_steps"]
    assert sd["radii_max"].dtype == torch.int16 and sd["area_sum"].dtype == torch.int32
    assert (sd["radius3d_min"] == 1).all() and (sd["weights_max"] == 0).all()
    c.area_sum += 3
    c.radius3d_max += 1
    c.reset(6)
    assert c.area_sum.shape == (6,) and (c.area_sum == 0).all()
    assert c.radius3d_max.shape == (4,) and (c.radius3d_max == 2).all()   # not reset
