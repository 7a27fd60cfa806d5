"""The configuration-2b training objective (0.8 L1 + 0.2 (1 - SSIM), LoG's
loss: LoG/render/renderer.py:260-273, LoG/render/loss.py:6-44), pinned by
closed forms and a direct window sum."""
import math

import torch

from filtergs_amd import workloads


def test_ssim_window():
    w = workloads.ssim_window()
    assert w.shape == (11, 11)
    assert abs(float(w.sum()) - 1.0) < 1e-6
    # separable Gaussian, sigma 1.5: w[5+i, 5+j] / w[5, 5] = exp(-(i^2 + j^2) / 4.5)
    for i, j in ((1, 0), (2, 3), (5, 5)):
        assert abs(float(w[5 + i, 5 + j] / w[5, 5]) - math.exp(-(i * i + j * j) / 4.5)) < 1e-6


def test_loss_closed_forms():
    g = torch.Generator().manual_seed(0)
    x = torch.rand((3, 30, 40), generator=g)
    assert float(workloads.l1_ssim_loss(x, x)) == 0.0
    # constant images a, b: every window has zero variance, SSIM =
    # (2ab + C1) / (a^2 + b^2 + C1), L1 = |a - b|
    a, b = 0.3, 0.7
    c1 = 1e-4
    want = 0.8 * (b - a) + 0.2 * (1 - (2 * a * b + c1) / (a * a + b * b + c1))
    f64 = dict(dtype=torch.float64)
    got = float(workloads.l1_ssim_loss(torch.full((3, 20, 20), a, **f64), torch.full((3, 20, 20), b, **f64),
                                       window=workloads.ssim_window().double() / workloads.ssim_window().double().sum()))
    assert abs(got - want) < 1e-12


def test_ssim_against_direct_window_sum():
    g = torch.Generator().manual_seed(1)
    x = torch.rand((1, 3, 16, 15), generator=g, dtype=torch.float64)
    y = torch.rand((1, 3, 16, 15), generator=g, dtype=torch.float64)
    w = workloads.ssim_window().double()
    # SSIM mean by explicit sums over every valid 11x11 window
    vals = []
    for c in range(3):
        for i in range(16 - 10):
            for j in range(15 - 10):
                px, py = x[0, c, i:i + 11, j:j + 11], y[0, c, i:i + 11, j:j + 11]
                mx, my = (w * px).sum(), (w * py).sum()
                vx = (w * px * px).sum() - mx * mx
                vy = (w * py * py).sum() - my * my
                cxy = (w * px * py).sum() - mx * my
                vals.append(((2 * mx * my + 1e-4) * (2 * cxy + 9e-4)) /
                            ((mx * mx + my * my + 1e-4) * (vx + vy + 9e-4)))
    want = 0.8 * float((x - y).abs().mean()) + 0.2 * (1 - float(torch.stack(vals).mean()))
    got = float(workloads.l1_ssim_loss(x, y, window=w))
    assert abs(got - want) < 1e-12
