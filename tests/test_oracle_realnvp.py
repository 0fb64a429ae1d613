"""Pins of the RealNVP oracle (oracle/realnvp.py); CPU only.

Checked by what the mathematics fixes: exact invertibility, log|det J| from a
full autograd Jacobian, normalisation of exp(log_prob) by 2-D quadrature,
the identity map at zero output weights, the masks of realNVP.py:155-158,
and the shapes of the reference's own tests (test_nf.py:24-52).
"""

import math

import torch

from oracle import realnvp as R


def params(d, L, h, seed=0):
    return R.init_params(d, L, h, seed=seed, scale=0.5, out_scale=0.4)


def test_masks_alternate_halves():
    assert R.mask(5, 0).tolist() == [1, 1, 0, 0, 0]
    assert R.mask(5, 1).tolist() == [0, 0, 1, 1, 1]
    assert R.mask(4, 2).tolist() == [1, 1, 0, 0]
    assert R.mask(1, 0).tolist() == [0] and R.mask(1, 1).tolist() == [1]


def test_inverse_and_log_det_signs():
    # test_nf.py:24-52: inverse(forward(x)) = x, log_det = -log_det_inverse
    p = params(3, 2, 4)
    x = torch.tensor([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]], dtype=torch.float64)
    y, ld = R.forward(p, x)
    xi, ldi = R.inverse(p, y)
    assert y.shape == x.shape and ld.shape == (2,)
    assert torch.allclose(xi, x, atol=1e-12) and torch.allclose(ld, -ldi, atol=1e-12)


def test_log_det_is_log_abs_det_jacobian():
    p = params(4, 3, 8, seed=1)
    x = torch.randn(3, 4, dtype=torch.float64)
    _, ld = R.forward(p, x)
    for i in range(3):
        J = torch.autograd.functional.jacobian(lambda v: R.forward(p, v[None])[0][0], x[i])
        assert abs(float(torch.linalg.slogdet(J)[1]) - float(ld[i])) < 1e-10


def test_density_integrates_to_one_in_2d():
    p = params(2, 4, 16, seed=2)
    p["data_mean"] = torch.tensor([0.3, -0.2], dtype=torch.float64)
    p["data_cov"] = torch.diag(torch.tensor([1.5, 0.7], dtype=torch.float64))
    g = torch.linspace(-12, 12, 801, dtype=torch.float64)
    X, Y = torch.meshgrid(g, g, indexing="ij")
    pts = torch.stack([X.reshape(-1), Y.reshape(-1)], -1)
    dens = torch.exp(R.log_prob(p, pts)).reshape(801, 801)
    h = float(g[1] - g[0])
    total = float(dens.sum()) * h * h
    # log_prob has no -sum log sigma term for the standardisation (realNVP.py:207-214):
    # the integral is prod sigma
    assert abs(total - math.sqrt(1.5 * 0.7)) < 1e-3, total


def test_zero_output_weights_give_the_identity():
    p = params(6, 3, 5)
    for lp in p["layers"]:
        for k in ("Ws2", "bs2", "Wt2", "bt2"):
            lp[k].zero_()
    x = torch.randn(10, 6, dtype=torch.float64)
    y, ld = R.forward(p, x)
    assert torch.equal(y, x) and torch.equal(ld, torch.zeros(10, dtype=torch.float64))


def test_sample_from_noise_inverts_log_prob_standardisation():
    p = params(3, 4, 6, seed=4)
    p["data_mean"] = torch.tensor([1.0, -2.0, 0.5], dtype=torch.float64)
    p["data_cov"] = torch.diag(torch.tensor([4.0, 0.25, 1.0], dtype=torch.float64))
    xi = torch.randn(20, 3, dtype=torch.float64)
    x = R.sample_from_noise(p, xi)
    z = (x - p["data_mean"]) / torch.sqrt(torch.diagonal(p["data_cov"]))
    y, _ = R.forward(p, z)
    assert torch.allclose(y, xi, atol=1e-10)


def test_flat_layout_round_trip():
    p = params(5, 3, 7)
    th = R.to_flat(p)
    assert th.numel() == 3 * 2 * (2 * 7 * 5 + 7 + 5) + 2 * 5 + 2 * 25
    q = R.from_flat(th, 5, 3, 7)
    x = torch.randn(4, 5, dtype=torch.float64)
    assert torch.equal(R.log_prob(p, x), R.log_prob(q, x))
