"""Built-in log-densities with hand-written device gradients.

A ``Target`` owns its parameter block on the device and exposes the C-ABI
descriptor consumed by the fused chain kernels. Wrapped in ``LogPDF`` it takes
the place of a user JAX function plus ``jax.value_and_grad`` in the reference
(src/flowMC/resource/logPDF.py:20-81, MALA.py:59, HMC.py:75-77).

``target_log_prob(target, x)`` is a ``torch.autograd.Function``: the forward
pass runs fmc_target_eval (value + gradient in one kernel) and the backward
pass returns the saved hand-written gradient.
"""

from __future__ import annotations

import math
from typing import Optional

import ctypes as C
import torch

from . import _native

KIND_GAUSS_DIAG = 0
KIND_GAUSS_DENSE = 1
KIND_ROSENBROCK = 2
KIND_FUNNEL = 3
KIND_GMM_DIAG = 4
KIND_DUAL_MOON = 5


def _f32(t, device) -> torch.Tensor:
    return torch.as_tensor(t, dtype=torch.float32).to(device).contiguous()


class Target:
    """Base class: a log-density evaluated on the device."""

    kind: int
    n_dim: int
    n_modes: int = 0

    def __init__(self, n_dim: int, params: torch.Tensor, device="cuda"):
        self.n_dim = int(n_dim)
        self.device = torch.device(device)
        self.params = params.to(self.device, torch.float32).contiguous().reshape(-1)
        if self.params.numel() == 0:
            self.params = torch.zeros(1, dtype=torch.float32, device=self.device)
            self._n_params = 0
        else:
            self._n_params = self.params.numel()

    def descriptor(self) -> _native.fmc_target:
        return _native.fmc_target(self.kind, self.n_dim, self.n_modes, 0,
                                  _native.dptr(self.params, name="target params"), self._n_params)

    # (value, gradient) at x [..., d]
    def value_and_grad(self, x: torch.Tensor, need_grad: bool = True):
        _native.require_gpu()
        lead = x.shape[:-1]
        if x.shape[-1] != self.n_dim:
            raise ValueError(f"expected last dim {self.n_dim}, got {x.shape[-1]}")
        xf = x.detach().to(torch.float32).reshape(-1, self.n_dim).contiguous()
        n = xf.shape[0]
        lp = torch.empty(n, dtype=torch.float32, device=xf.device)
        g = torch.empty_like(xf) if need_grad else None
        desc = self.descriptor()
        _native.call("fmc_target_eval", _native.stream_ptr(), C.byref(desc), _native.dptr(xf, name="x"), n,
                     _native.dptr(lp, name="logp"), _native.dptr(g, name="grad"), None)
        lp = lp.reshape(lead)
        return (lp, g.reshape(*lead, self.n_dim)) if need_grad else (lp, None)

    def __call__(self, x: torch.Tensor, data=None) -> torch.Tensor:
        return target_log_prob(self, x)

    def __repr__(self):
        return f"{type(self).__name__}(n_dim={self.n_dim})"


class GaussianDiag(Target):
    """log p = -1/2 sum_j prec_j (x_j - mean_j)^2 (unnormalised)."""

    kind = KIND_GAUSS_DIAG

    def __init__(self, mean, prec=None, device="cuda"):
        mean = torch.as_tensor(mean, dtype=torch.float32).reshape(-1)
        prec = torch.ones_like(mean) if prec is None else torch.as_tensor(prec, dtype=torch.float32).reshape(-1)
        self.mean, self.prec = mean, prec
        super().__init__(mean.numel(), torch.cat([mean, prec]), device)


class GaussianDense(Target):
    """log p = -1/2 (x - mean)^T P (x - mean) with P the precision matrix (symmetrised)."""

    kind = KIND_GAUSS_DENSE

    def __init__(self, mean, precision, device="cuda"):
        mean = torch.as_tensor(mean, dtype=torch.float32).reshape(-1)
        P = torch.as_tensor(precision, dtype=torch.float64)
        P = (0.5 * (P + P.T)).to(torch.float32)
        self.mean, self.precision = mean, P
        super().__init__(mean.numel(), torch.cat([mean, P.reshape(-1)]), device)

    @classmethod
    def from_covariance(cls, mean, cov, device="cuda"):
        return cls(mean, torch.linalg.inv(torch.as_tensor(cov, dtype=torch.float64)), device)


class Rosenbrock(Target):
    """log p = -scale * sum_i [100 (x_{i+1} - x_i^2)^2 + (1 - x_i)^2], scale = 1/20 by default."""

    kind = KIND_ROSENBROCK

    def __init__(self, n_dim: int, scale: float = 1.0 / 20.0, device="cuda"):
        self.scale = float(scale)
        super().__init__(n_dim, torch.tensor([self.scale]), device)


class Funnel(Target):
    """Neal's funnel: x_0 = v ~ N(0, 3^2), x_i | v ~ N(0, e^v) (unnormalised)."""

    kind = KIND_FUNNEL

    def __init__(self, n_dim: int, device="cuda"):
        if n_dim < 2:
            raise ValueError("funnel needs n_dim >= 2")
        super().__init__(n_dim, torch.zeros(0, dtype=torch.float32), device)


class GaussianMixtureDiag(Target):
    """log p = logsumexp_k [log w_k + sum_j log N(x_j; mu_kj, s_kj^2)] + d/2 log(2 pi)."""

    kind = KIND_GMM_DIAG

    def __init__(self, means, sigmas, weights=None, device="cuda"):
        means = torch.as_tensor(means, dtype=torch.float64)
        sigmas = torch.as_tensor(sigmas, dtype=torch.float64)
        K, d = means.shape
        w = torch.full((K,), 1.0 / K, dtype=torch.float64) if weights is None else torch.as_tensor(
            weights, dtype=torch.float64)
        c = torch.log(w) - torch.log(sigmas).sum(-1)
        self.means, self.sigmas, self.weights = means, sigmas, w
        self.n_modes = int(K)
        super().__init__(d, torch.cat([means.reshape(-1), (1.0 / sigmas).reshape(-1), c]).to(torch.float32), device)

    def mean(self) -> torch.Tensor:
        return (self.weights[:, None] * self.means).sum(0)

    def variance(self) -> torch.Tensor:
        m = self.mean()
        return (self.weights[:, None] * (self.sigmas ** 2 + self.means ** 2)).sum(0) - m ** 2


class DualMoon(Target):
    """Dual-moon target of docs/tutorials/dualmoon.ipynb (center 0) and
    test/integration/test_MALA.py:11-20 (center = data)."""

    kind = KIND_DUAL_MOON

    def __init__(self, n_dim: int, center=None, device="cuda"):
        if n_dim < 2:
            raise ValueError("dual moon needs n_dim >= 2")
        c = torch.zeros(n_dim, dtype=torch.float32) if center is None else torch.as_tensor(center, dtype=torch.float32).reshape(-1)
        self.center = c
        super().__init__(n_dim, c, device)


class _TargetLogProb(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, target):
        lp, g = target.value_and_grad(x, need_grad=x.requires_grad)
        if g is not None:
            ctx.save_for_backward(g)
        return lp.to(x.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return (grad_out[..., None] * g).to(grad_out.dtype), None


def target_log_prob(target: Target, x: torch.Tensor) -> torch.Tensor:
    """Differentiable log-density: backward uses the kernel's hand-written gradient."""
    return _TargetLogProb.apply(x, target)


def gaussian_mixture_64d(seed: int = 0, n_dim: int = 64, n_modes: int = 8, device="cuda"):
    """The 8-mode mixture of the full NF-MCMC workload: modes at seeded +-4
    vertices, per-mode diagonal sigma ~ U(0.5, 1.5), equal weights."""
    g = torch.Generator().manual_seed(seed)
    signs = torch.randint(0, 2, (n_modes, n_dim), generator=g).to(torch.float64) * 2 - 1
    means = 4.0 * signs
    sigmas = 0.5 + torch.rand((n_modes, n_dim), generator=g, dtype=torch.float64)
    return GaussianMixtureDiag(means, sigmas, device=device)


def correlated_gaussian_5d(seed: int = 0, n_dim: int = 5, device="cuda"):
    """Sigma = Q diag(lam) Q^T, Q seeded orthogonal, lam log-uniform in [0.1, 10]."""
    g = torch.Generator().manual_seed(seed)
    A = torch.randn((n_dim, n_dim), generator=g, dtype=torch.float64)
    Q, R = torch.linalg.qr(A)
    Q = Q * torch.sign(torch.diagonal(R))[None, :]
    lam = 10.0 ** (torch.rand(n_dim, generator=g, dtype=torch.float64) * 2 - 1)
    cov = Q @ torch.diag(lam) @ Q.T
    t = GaussianDense.from_covariance(torch.zeros(n_dim, dtype=torch.float32), cov, device=device)
    t.covariance = cov
    return t


def funnel_dims(seed: int = 0, n_targets: int = 32, lo: int = 8, hi: int = 256) -> list[int]:
    """Target dimensions drawn log-uniformly in [lo, hi]."""
    g = torch.Generator().manual_seed(seed)
    u = torch.rand(n_targets, generator=g, dtype=torch.float64)
    return [int(round(math.exp(math.log(lo) + float(v) * (math.log(hi) - math.log(lo))))) for v in u]
