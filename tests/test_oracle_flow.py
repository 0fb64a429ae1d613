"""Pins the oracle flow / spline / optimizer to facts fixed by the mathematics
or by independent code (torch.autograd.functional.jacobian, torch.optim.AdamW)."""

import math

import pytest
import torch

from oracle import flow as F
from oracle import optim as O


@pytest.fixture(autouse=True, scope="module")
def _float64_default():
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


def rand_raw(n, K, seed, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(n, 3 * K + 1, generator=g) * scale


def test_spline_identity_at_zero_logits():
    K = 8
    raw = torch.zeros(50, 3 * K + 1)
    xp, yp, s = F.spline_params(raw, K)
    assert torch.allclose(s, torch.ones_like(s), atol=1e-12)
    x = torch.linspace(-14, 14, 50)
    y, ld = F.spline_fwd(x, xp, yp, s)
    assert torch.allclose(y, x, atol=1e-12) and torch.allclose(ld, torch.zeros_like(ld), atol=1e-12)


def test_spline_knot_values_and_slopes():
    # y(x_k) = y_k and dy/dx(x_k) = d_k: the defining interpolation property of the RQ spline
    K = 5
    raw = rand_raw(1, K, 1)
    xp, yp, s = F.spline_params(raw, K)
    xp, yp, s = xp[0], yp[0], s[0]
    for k in range(1, K):
        y, _ = F.spline_fwd(xp[k:k + 1], xp[None], yp[None], s[None])
        assert abs(float(y) - float(yp[k])) < 1e-12
    h = 1e-7
    for k in range(K + 1):
        x0 = float(xp[k]) + (h if k < K else -h) * 10
        xs = torch.tensor([x0 - h, x0 + h])
        y, _ = F.spline_fwd(xs, xp.expand(2, -1), yp.expand(2, -1), s.expand(2, -1))
        assert abs(float((y[1] - y[0]) / (2 * h)) - float(s[k])) < 1e-4


def test_spline_inverse_and_logdet_by_fd():
    K = 8
    raw = rand_raw(400, K, 2)
    xp, yp, s = F.spline_params(raw, K)
    x = torch.linspace(-12, 12, 400)
    y, ld = F.spline_fwd(x, xp, yp, s)
    xr, ldi = F.spline_inv(y, xp, yp, s)
    assert torch.allclose(xr, x, atol=1e-9)
    assert torch.allclose(ld, -ldi, atol=1e-9)
    h = 1e-6
    y1, _ = F.spline_fwd(x + h, xp, yp, s)
    y0, _ = F.spline_fwd(x - h, xp, yp, s)
    assert torch.allclose(torch.log((y1 - y0) / (2 * h)), ld, atol=1e-5)
    assert bool(((y1 - y0) > 0).all())  # monotone


def test_flow_change_of_variables_full_jacobian():
    d, L, K = 3, 3, 4
    p = F.init_params(d, L, K, 8, 8, seed=3, scale=0.3)
    x = torch.randn(5, d) * 2
    y, ld = F.forward(p, x, K)
    for i in range(5):
        J = torch.autograd.functional.jacobian(lambda z: F.forward(p, z[None], K)[0][0], x[i])
        assert abs(float(torch.linalg.slogdet(J)[1]) - float(ld[i])) < 1e-9
    # masked (conditioning) dims pass through a coupling unchanged
    yc, _ = F.coupling(p["layers"][0], x, 0, K, False)
    m = F.mask(d, 0).bool()
    assert torch.equal(yc[:, m], x[:, m])


def test_flow_density_integrates_to_one_2d():
    d, L, K = 2, 4, 6
    p = F.init_params(d, L, K, 8, 8, seed=4, scale=0.1)
    g = torch.linspace(-14, 14, 561)
    X, Y = torch.meshgrid(g, g, indexing="ij")
    pts = torch.stack([X.reshape(-1), Y.reshape(-1)], -1)
    lp = F.log_prob(p, pts, K).reshape(561, 561)
    integral = torch.trapezoid(torch.trapezoid(torch.exp(lp), g, dim=1), g)
    assert abs(float(integral) - 1.0) < 2e-3


def test_reference_inverse_order_quirk():
    d, L, K = 4, 2, 4
    p = F.init_params(d, L, K, 8, 8, seed=5, scale=0.5)
    x = torch.randn(6, d)
    y, _ = F.forward(p, x, K)
    xe, lde = F.inverse(p, y, K, exact=True)
    assert torch.allclose(xe, x, atol=1e-9)
    xr, _ = F.inverse(p, y, K, exact=False)
    assert not torch.allclose(xr, x, atol=1e-3)  # reference order is not the inverse once affine != 0


def test_loss_gradient_matches_finite_differences():
    d, L, K = 3, 2, 3
    p = F.init_params(d, L, K, 6, 6, seed=6, scale=0.2)
    theta = F.to_flat(p).clone().requires_grad_(True)
    x = torch.randn(7, d)
    loss = F.loss(F.from_flat(theta, d, L, K, 6, 6), x, K)
    (g,) = torch.autograd.grad(loss, theta)
    g2 = torch.Generator().manual_seed(0)
    n_layer = F.to_flat({"layers": p["layers"][:1], "data_mean": torch.zeros(0), "data_cov": torch.zeros(0),
                         "base_mean": torch.zeros(0), "base_cov": torch.zeros(0)}).numel()
    idx = torch.randint(0, n_layer * L, (25,), generator=g2)
    h = 1e-6
    for i in idx.tolist():
        tp = theta.detach().clone()
        tm = theta.detach().clone()
        tp[i] += h
        tm[i] -= h
        fd = (F.loss(F.from_flat(tp, d, L, K, 6, 6), x, K) - F.loss(F.from_flat(tm, d, L, K, 6, 6), x, K)) / (2 * h)
        assert abs(float(fd) - float(g[i])) < 1e-6 * (1 + abs(float(g[i])))


@pytest.mark.parametrize("clip", [True, False])
def test_adamw_clip_matches_torch_optim(clip):
    g = torch.Generator().manual_seed(7)
    theta = torch.randn(50, generator=g)
    ref = theta.clone().requires_grad_(True)
    opt = torch.optim.AdamW([ref], lr=1e-2, betas=(0.9, 0.999), eps=1e-8, weight_decay=1e-4)
    m = torch.zeros(50)
    v = torch.zeros(50)
    th = theta.clone()
    for t in range(1, 4):
        grad = torch.randn(50, generator=g) * (3.0 if clip else 0.01)
        th, m, v = O.adamw_clip_step(th, grad, m, v, t, 1e-2, max_norm=1.0)
        ref.grad = grad.clone()
        # torch's clip uses max_norm / (norm + 1e-6): equal to optax's rule up to ~1e-6 relative
        torch.nn.utils.clip_grad_norm_([ref], max_norm=1.0)
        opt.step()
    assert torch.allclose(th, ref.detach(), atol=1e-12 if not clip else 1e-7)
