"""Flow building blocks — src/flowMC/resource/model/common.py.

The device flow keeps all parameters in one flat FP32 vector; the classes
here are views onto it that expose the reference's attribute layout
(``MLP.layers``, ``ScalarAffine.scale/shift``, ``Gaussian.mean/cov``,
``MaskedCouplingLayer.mask``) for inspection and editing.
"""

from __future__ import annotations

from abc import ABC, abstractmethod

import math

import torch


class Bijection(ABC):
    """common.py:10-41."""

    def __call__(self, x, condition=None):
        return self.forward(x, condition)

    @abstractmethod
    def forward(self, x, condition=None):
        raise NotImplementedError

    @abstractmethod
    def inverse(self, x, condition=None):
        raise NotImplementedError


class Distribution(ABC):
    """common.py:44-65."""

    def __call__(self, x, key=None):
        return self.log_prob(x)

    @abstractmethod
    def log_prob(self, x):
        raise NotImplementedError

    @abstractmethod
    def sample(self, rng_key, n_samples: int):
        raise NotImplementedError


class Linear:
    """View of one eqx.nn.Linear (y = W x + b)."""

    def __init__(self, weight: torch.Tensor, bias: torch.Tensor):
        self.weight = weight
        self.bias = bias

    @property
    def in_features(self):
        return self.weight.shape[1]

    @property
    def out_features(self):
        return self.weight.shape[0]

    def __call__(self, x):
        return x @ self.weight.T + self.bias


class MLP:
    """View of the conditioner (common.py:68-124): Linear and tanh alternating,
    one Linear per hidden layer, then the output Linear."""

    def __init__(self, linears: list):
        self.layers = []
        for i, lin in enumerate(linears):
            self.layers.append(lin)
            if i < len(linears) - 1:
                self.layers.append(torch.tanh)

    def __call__(self, x):
        for layer in self.layers:
            x = layer(x)
        return x

    @property
    def n_input(self) -> int:
        return self.layers[0].in_features

    @property
    def n_output(self) -> int:
        return self.layers[-1].out_features

    @property
    def dtype(self):
        return self.layers[0].weight.dtype


class ScalarAffine(Bijection):
    """View of ScalarAffine (common.py:211-240): y = (x + shift) e^scale."""

    def __init__(self, scale: torch.Tensor, shift: torch.Tensor):
        self.scale = scale
        self.shift = shift

    def forward(self, x, condition=None):
        return (x + self.shift) * torch.exp(self.scale), self.scale

    def inverse(self, x, condition=None):
        return x * torch.exp(-self.scale) - self.shift, -self.scale


class MaskedCouplingLayer(Bijection):
    """View of MaskedCouplingLayer (common.py:127-168); evaluation goes through the flow kernels."""

    def __init__(self, bijector, mask: torch.Tensor):
        self.bijector = bijector
        self._mask = mask

    @property
    def mask(self):
        return self._mask

    def forward(self, x, condition=None):
        raise NotImplementedError("individual layers are evaluated by the fused flow kernels")

    def inverse(self, x, condition=None):
        raise NotImplementedError("individual layers are evaluated by the fused flow kernels")


class Gaussian(Distribution):
    """Multivariate Gaussian base distribution (common.py:243-293). The flow
    kernels evaluate it inside their passes (diagonal fast path, or the full
    covariance through its Cholesky factor); these methods are the same
    densities for direct use."""

    def __init__(self, mean: torch.Tensor, cov: torch.Tensor, learnable: bool = False):
        self._mean = torch.as_tensor(mean, dtype=torch.float32)
        self._cov = torch.as_tensor(cov, dtype=torch.float32)
        self.learnable = learnable

    @property
    def mean(self):
        return self._mean

    @property
    def cov(self):
        return self._cov

    def log_prob(self, x):
        """multivariate_normal.logpdf(x, mean, cov) (common.py:285-286)."""
        d = self._mean.shape[-1]
        L = torch.linalg.cholesky(self._cov.double())
        u = torch.as_tensor(x, dtype=torch.float64, device=L.device) - self._mean.to(L.device, torch.float64)
        w = torch.linalg.solve_triangular(L, u.reshape(-1, d).T, upper=False).T.reshape(u.shape)
        lp = -0.5 * (w * w).sum(-1) - torch.log(torch.diagonal(L)).sum() - 0.5 * d * math.log(2 * math.pi)
        return lp.to(torch.float32)

    def sample(self, rng_key, n_samples: int):
        """mean + xi @ chol(cov).T (jax.random.multivariate_normal, common.py:288-293)."""
        from ... import random as fr

        z = fr.normal(rng_key, (n_samples, self._mean.shape[-1]), device=self._cov.device)
        L = torch.linalg.cholesky(self._cov.double()).to(torch.float32)
        return self._mean.to(z.device) + z @ L.T.to(z.device)
