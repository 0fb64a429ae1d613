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


@pytest.mark.parametrize("hidden", [[8], [8, 8], [8, 6, 8], [8, 5, 7, 6]])
def test_mlp_any_depth_matches_torch_sequential(hidden):
    """The conditioner of n hidden layers equals torch.nn's Linear / Tanh
    stack built from the same weights (common.py:86-112): every hidden layer
    applied once, tanh after each hidden layer, none after the output."""
    d, K = 4, 3
    lp = F.init_params(d, 1, K, hidden, seed=11, hidden_var=2.0)["layers"][0]
    widths = [d] + list(hidden) + [d * (3 * K + 1)]
    mods = []
    for i in range(len(widths) - 1):
        lin = torch.nn.Linear(widths[i], widths[i + 1], dtype=torch.float64)
        with torch.no_grad():
            lin.weight.copy_(lp[f"W{i + 1}"])
            lin.bias.copy_(lp[f"b{i + 1}"])
        mods.append(lin)
        if i < len(widths) - 2:
            mods.append(torch.nn.Tanh())
    net = torch.nn.Sequential(*mods)
    c = torch.randn(9, d) * 2
    with torch.no_grad():
        assert torch.allclose(F.mlp(lp, c), net(c), atol=1e-12)
        # a conditioner that skips its last hidden layer (or puts tanh on the output) differs
        bad = dict(lp)
        assert F.n_linear(lp) == len(hidden) + 1
        if len(hidden) > 1:
            n = len(hidden)
            bad.pop(f"W{n + 1}"), bad.pop(f"b{n + 1}")
            bad[f"W{n}"] = torch.randn(widths[-1], widths[n - 1])
            bad[f"b{n}"] = lp[f"b{n + 1}"]
            assert not torch.allclose(F.mlp(bad, c), net(c), atol=1e-3)


@pytest.mark.parametrize("hidden", [[8], [8, 8], [8, 6, 8, 5]])
def test_flow_change_of_variables_full_jacobian(hidden):
    d, L, K = 3, 3, 4
    p = F.init_params(d, L, K, hidden, seed=3, scale=0.3)
    x = torch.randn(5, d) * 2
    y, ld = F.forward(p, x, K)
    for i in range(5):
        J = torch.autograd.functional.jacobian(lambda z: F.forward(p, z[None], K)[0][0], x[i])
        assert abs(float(torch.linalg.slogdet(J)[1]) - float(ld[i])) < 1e-9
    # masked (conditioning) dims pass through a coupling unchanged
    yc, _ = F.coupling(p["layers"][0], x, 0, K, False)
    m = F.mask(d, 0).bool()
    assert torch.equal(yc[:, m], x[:, m])


def _dense_base(p, d):
    """A correlated base Gaussian (full covariance, nonzero mean)."""
    g = torch.Generator().manual_seed(12)
    A = torch.randn(d, d, generator=g)
    p["base_cov"] = A @ A.T / d + 0.5 * torch.eye(d)
    p["base_mean"] = torch.randn(d, generator=g) * 0.5
    return p


@pytest.mark.parametrize("dense_base", [False, True])
def test_flow_density_integrates_to_one_2d(dense_base):
    d, L, K = 2, 4, 6
    p = F.init_params(d, L, K, 8, 8, seed=4, scale=0.1)
    if dense_base:
        p = _dense_base(p, d)
    g = torch.linspace(-14, 14, 561)
    X, Y = torch.meshgrid(g, g, indexing="ij")
    pts = torch.stack([X.reshape(-1), Y.reshape(-1)], -1)
    lp = F.log_prob(p, pts, K).reshape(561, 561)
    integral = torch.trapezoid(torch.trapezoid(torch.exp(lp), g, dim=1), g)
    assert abs(float(integral) - 1.0) < 2e-3


def test_identity_flow_with_dense_base_is_the_base_gaussian():
    """With zero conditioner output weights and zero ScalarAffine parameters
    every layer is the identity (rqSpline.py:20-39: zero logits give the
    identity spline), so log_prob is the base MVN density (scipy) and the
    samples have the base mean and covariance (jax.random.multivariate_normal
    draws mean + chol(cov) xi, common.py:276-293)."""
    from scipy.stats import multivariate_normal

    d, L, K = 5, 2, 4
    p = _dense_base(F.init_params(d, L, K, [6, 6], seed=13), d)
    for lp in p["layers"]:
        n = F.n_linear(lp)
        lp[f"W{n}"].zero_()
        lp[f"b{n}"].zero_()
    x = torch.randn(40, d) * 1.5
    ref = multivariate_normal(p["base_mean"].numpy(), p["base_cov"].numpy()).logpdf(x.numpy())
    assert torch.allclose(F.log_prob(p, x, K), torch.as_tensor(ref), atol=1e-10)
    xi = torch.randn(200000, d, generator=torch.Generator().manual_seed(14))
    z = F.sample_from_noise(p, xi, K, exact=True)
    assert torch.allclose(z.mean(0), p["base_mean"], atol=0.02)
    assert torch.allclose(torch.cov(z.T), p["base_cov"], atol=0.03)


def test_learnable_base_gradient_is_the_gaussian_closed_form():
    """Identity flow (zero output layers and affines), learnable base
    (common.py:243-273): the autograd gradient of sum_n w_n log_prob(x_n)
    with respect to the base mean is sum_n w_n S^-1 r_n and with respect to
    the covariance entries 1/2 sum_n w_n (S^-1 r_n r_n^T S^-1 - S^-1), the
    symmetric matrix that multivariate_normal.logpdf's symmetrised Cholesky
    input gives (r_n = x_n - mean). A dropped 1/2, a dropped -S^-1 term or a
    sign error fails it; a frozen base gets no gradient."""
    d, L, K = 4, 2, 5
    p = _dense_base(F.init_params(d, L, K, [6, 6], seed=21), d)
    for lp in p["layers"]:
        n = F.n_linear(lp)
        lp[f"W{n}"].zero_()
        lp[f"b{n}"].zero_()
    g = torch.Generator().manual_seed(22)
    x = torch.randn(30, d, generator=g) * 1.3
    w = torch.randn(30, generator=g)
    mu = p["base_mean"].clone().requires_grad_(True)
    cov = p["base_cov"].clone().requires_grad_(True)
    q = dict(p, base_mean=mu, base_cov=cov)
    gm, gc = torch.autograd.grad((w * F.log_prob(q, x, K, learnable_base=True)).sum(), [mu, cov])
    S = p["base_cov"]
    Si = torch.linalg.inv(S)
    r = (x - p["base_mean"]) @ Si  # rows S^-1 r_n (S symmetric)
    want_m = (w[:, None] * r).sum(0)
    want_c = 0.5 * (r.T @ (w[:, None] * r)) - 0.5 * w.sum() * Si
    assert torch.allclose(gm, want_m, atol=1e-10)
    assert torch.allclose(gc, want_c, atol=1e-10)
    assert torch.allclose(gc, gc.T, atol=1e-12)
    frozen = dict(p, base_mean=p["base_mean"].clone().requires_grad_(True),
                  base_cov=p["base_cov"].clone().requires_grad_(True))
    assert not F.log_prob(frozen, x, K).requires_grad  # the only leaves needing a gradient are detached


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
