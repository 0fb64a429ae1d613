"""FP64 log-densities of the built-in targets (value only; gradients by autograd).

The gradient the kernels hand-code is checked against torch.autograd of these
values (the role jax.value_and_grad plays at src/flowMC/resource/kernel/MALA.py:59),
and autograd itself against central finite differences in the tests.
Each function maps x [..., d] -> [...].
"""

from __future__ import annotations

import math

import torch


def gauss_diag(x, mean, prec):
    y = x - mean
    return -0.5 * (prec * y * y).sum(-1)


def gauss_dense(x, mean, P):
    y = x - mean
    return -0.5 * torch.einsum("...i,ij,...j->...", y, P, y)


def rosenbrock(x, scale=1.0 / 20.0):
    a = x[..., 1:] - x[..., :-1] ** 2
    b = 1.0 - x[..., :-1]
    return -scale * (100.0 * a * a + b * b).sum(-1)


def funnel(x):
    v = x[..., 0]
    rest = x[..., 1:]
    d1 = x.shape[-1] - 1
    return -v * v / 18.0 - 0.5 * d1 * v - 0.5 * torch.exp(-v) * (rest * rest).sum(-1)


def gmm_diag(x, means, sigmas, log_w=None):
    """means, sigmas [K, d]; equal weights unless log_w [K] is given."""
    K = means.shape[0]
    if log_w is None:
        log_w = torch.full((K,), -math.log(K), dtype=x.dtype)
    z = (x[..., None, :] - means) / sigmas  # [..., K, d]
    q = log_w - torch.log(sigmas).sum(-1) - 0.5 * (z * z).sum(-1)
    return torch.logsumexp(q, dim=-1)


def dual_moon(x, center=None):
    """docs/tutorials/dualmoon.ipynb (center=0) / test/integration/test_MALA.py:11-20."""
    y = x if center is None else x - center
    n = torch.linalg.norm(y, dim=-1)
    term1 = 0.5 * ((n - 2.0) / 0.1) ** 2
    s = torch.tensor([-3.0, 3.0], dtype=x.dtype)
    term2 = -0.5 * ((x[..., :1] + s) / 0.8) ** 2
    term3 = -0.5 * ((x[..., 1:2] + s) / 0.6) ** 2
    return -(term1 - torch.logsumexp(term2, -1) - torch.logsumexp(term3, -1))


def value_and_grad(fn, x):
    """Batched value and gradient of a per-point log-density via autograd (FP64)."""
    x = x.detach().clone().requires_grad_(True)
    v = fn(x)
    (g,) = torch.autograd.grad(v.sum(), x)
    return v.detach(), g.detach()
