"""Pins the oracle targets: closed forms, known answer, autograd vs finite differences."""

import math

import pytest
import torch

from oracle import targets as T



@pytest.fixture(autouse=True, scope="module")
def _float64_default():
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


def fd_grad(fn, x, h=1e-6):
    g = torch.zeros_like(x)
    for i in range(x.shape[-1]):
        e = torch.zeros_like(x)
        e[..., i] = h
        g[..., i] = (fn(x + e) - fn(x - e)) / (2 * h)
    return g


def test_dual_moon_known_answer():
    # docs/tutorials/dualmoon.ipynb:65 prints -218.14496 at x = 0 (d = 5)
    v = T.dual_moon(torch.zeros(5))
    assert abs(float(v) - (-218.14496)) < 5e-5


def test_rosenbrock_closed_form():
    assert float(T.rosenbrock(torch.ones(128))) == 0.0
    x = torch.zeros(3)  # terms: (1-0)^2 twice, 100*(0-0)^2
    assert abs(float(T.rosenbrock(x, scale=1.0)) + 2.0) < 1e-15


def test_gmm_single_mode_is_gaussian():
    mu = torch.tensor([[1.0, -2.0, 0.5]])
    s = torch.tensor([[0.5, 2.0, 1.5]])
    x = torch.randn(10, 3)
    want = torch.distributions.Normal(mu[0], s[0]).log_prob(x).sum(-1) + 1.5 * math.log(2 * math.pi)
    assert torch.allclose(T.gmm_diag(x, mu, s), want, atol=1e-12)


def test_funnel_matches_conditional_normals():
    x = torch.randn(20, 6)
    v = x[:, 0]
    want = torch.distributions.Normal(0.0, 3.0).log_prob(v) + torch.distributions.Normal(
        0.0, torch.exp(0.5 * v)[:, None]).log_prob(x[:, 1:]).sum(-1)
    const = -0.5 * math.log(2 * math.pi) * 6 - math.log(3.0)
    assert torch.allclose(T.funnel(x) + const, want, atol=1e-10)


def test_gauss_dense_matches_mvn():
    A = torch.randn(5, 5)
    cov = A @ A.T + 5 * torch.eye(5)
    P = torch.linalg.inv(cov)
    mu = torch.randn(5)
    x = torch.randn(7, 5)
    mvn = torch.distributions.MultivariateNormal(mu, cov)
    want = mvn.log_prob(x) - mvn.log_prob(mu)
    assert torch.allclose(T.gauss_dense(x, mu, P), want, atol=1e-10)


@pytest.mark.parametrize("name", ["gauss_diag", "rosenbrock", "funnel", "gmm", "dual_moon", "gauss_dense"])
def test_autograd_gradient_matches_finite_differences(name):
    g = torch.Generator().manual_seed(3)
    d = 6
    x = torch.randn(4, d, generator=g) * 0.7
    fns = {
        "gauss_diag": lambda z: T.gauss_diag(z, torch.arange(d) * 0.1, torch.linspace(0.5, 2, d)),
        "rosenbrock": lambda z: T.rosenbrock(z),
        "funnel": lambda z: T.funnel(z),
        "gmm": lambda z: T.gmm_diag(z, torch.randn(3, d, generator=torch.Generator().manual_seed(1)),
                                    0.5 + torch.rand(3, d, generator=torch.Generator().manual_seed(2))),
        "dual_moon": lambda z: T.dual_moon(z, torch.arange(d) * 0.1),
        "gauss_dense": lambda z: T.gauss_dense(z, torch.zeros(d), torch.eye(d) + 0.1),
    }
    fn = fns[name]
    _, ga = T.value_and_grad(fn, x)
    assert torch.allclose(ga, fd_grad(fn, x), rtol=1e-6, atol=1e-6)


def _mixture_reference(x, means, sigmas, log_w):
    # torch.distributions: categorical over modes of independent normals
    mix = torch.distributions.MixtureSameFamily(
        torch.distributions.Categorical(logits=log_w),
        torch.distributions.Independent(torch.distributions.Normal(means, sigmas), 1))
    return mix.log_prob(x) + 0.5 * x.shape[-1] * math.log(2 * math.pi)


def test_gmm_several_modes_matches_mixture_same_family():
    # K = 5 modes, unequal per-mode / per-dim sigmas, non-uniform weights: pins
    # the per-mode -sum log sigma_k and the log w_k the K = 1 case cannot see
    g = torch.Generator().manual_seed(7)
    K, d = 5, 6
    means = 3 * torch.randn(K, d, generator=g)
    sigmas = 0.4 + 1.5 * torch.rand(K, d, generator=g)
    w = torch.tensor([0.05, 0.4, 0.1, 0.3, 0.15])
    log_w = torch.log(w)
    x = means[torch.randint(0, K, (400,), generator=g)] + 1.3 * torch.randn(400, d, generator=g)
    want = _mixture_reference(x, means, sigmas, log_w)
    assert torch.allclose(T.gmm_diag(x, means, sigmas, log_w), want, atol=1e-10, rtol=1e-12)

    # planted mistakes (restated on the CPU) must fail the same comparison
    def dropped_log_sigma(z):
        zz = (z[..., None, :] - means) / sigmas
        return torch.logsumexp(log_w - 0.5 * (zz * zz).sum(-1), dim=-1)

    def uniform_weights(z):
        return T.gmm_diag(z, means, sigmas)

    for bad in (dropped_log_sigma, uniform_weights):
        assert not torch.allclose(bad(x), want, atol=1e-3)
