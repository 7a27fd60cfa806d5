"""Pins the oracle's training-side restatements: the Adam step
(LoG/model/sparse_optimizer.py:41-78) against closed forms -- the first step
moves a row by lr * sign(g), constant gradients converge to lr * sign(g) per
step, unselected rows do not move, amsgrad keeps the running maximum -- and
the counter update (LoG/model/counter.py:36-83) against hand-counted cases."""
import numpy as np

from oracle import pyoracle as orc


def _adam_inputs(N=50, D=3, seed=0):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(N, D)).astype(np.float32)
    z = np.zeros((N, D), np.float32)
    index = rng.choice(N, 20, replace=False)
    flag = rng.random(20) < 0.7
    g = rng.normal(size=(20, D)).astype(np.float32)
    return p, z, index, flag, g


def test_adam_first_step_is_sign_step():
    p, z, index, flag, g = _adam_inputs()
    lr = 1e-2
    p1, m, v, _ = orc.sparse_adam_step(p, p[index], g, z, z, index, flag, lr, step=1)
    sel = index[flag]
    np.testing.assert_allclose(p1[sel] - p[sel], -lr * np.sign(g[flag]), rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(m[sel], 0.1 * g[flag], rtol=1e-6)
    np.testing.assert_allclose(v[sel], 0.001 * g[flag] ** 2, rtol=1e-5)
    untouched = np.setdiff1d(np.arange(len(p)), sel)
    assert np.array_equal(p1[untouched], p[untouched]) and not m[untouched].any()


def test_adam_constant_gradient_and_amsgrad():
    p, z, index, flag, g = _adam_inputs(seed=3)
    lr = 5e-3
    m, v, vmax, cur = z, z, z, p
    for step in range(1, 40):
        prev = cur
        cur, m, v, vmax = orc.sparse_adam_step(cur, cur[index], g, m, v, index, flag, lr, step,
                                               max_exp_avg_sq=vmax)
    sel = index[flag]
    # with a constant gradient the bias-corrected ratio m / sqrt(v) is sign(g)
    np.testing.assert_allclose(cur[sel] - prev[sel], -lr * np.sign(g[flag]), rtol=1e-4, atol=1e-7)
    assert (vmax[sel] >= v[sel]).all()


def test_counter_update_hand_case():
    N = 8
    c = dict(create_steps=np.zeros(N, np.int32), visible_count=np.zeros(N, np.int16),
             weights_max=np.full(N, 0.25, np.float32), weights_sum=np.zeros(N, np.float32),
             radii_max=np.full(N, 3, np.int16), radii_max_max=np.full(N, 4, np.int32),
             area_sum=np.zeros(N, np.int32), grad_sum=np.zeros(N, np.float32))
    visible_index = np.array([5, 2, 7])
    radii = np.array([2, 0, 9], np.int32)
    w = np.array([0.5, 0.9, 0.1], np.float32)
    grad = np.array([[3.0, 4.0, 0.0], [0.0, 1.0, 0.0], [1.0, 0.0, 9.0]], np.float32)
    out = orc.counter_update(c, visible_index, radii, w, np.array([0, 2]), np.array([10, 2]), grad)
    assert list(out["create_steps"][[5, 2, 7]]) == [1, 0, 1]
    assert list(out["visible_count"][[5, 2, 7]]) == [1, 0, 1]
    assert list(out["weights_max"][[5, 2, 7]]) == [0.5, 0.25, 0.25]
    assert list(out["weights_sum"][[5, 2, 7]]) == [0.5, 0.0, np.float32(0.1)]
    assert list(out["radii_max"][[5, 2, 7]]) == [3, 3, 9]
    assert list(out["area_sum"][[5, 2, 7]]) == [10, 0, 2]
    assert list(out["grad_sum"][[5, 2, 7]]) == [50.0, 0.0, 2.0]
    assert list(out["radii_max_max"][[5, 2, 7]]) == [10, 4, 4]
