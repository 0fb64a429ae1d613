"""GPU parity of the RealNVP kernels against the FP64 oracle (oracle/realnvp.py).

Tolerances (FP32 kernels vs FP64 oracle on the FP32-rounded parameters):
forward / inverse / log_prob / samples within 2e-4 * (1 + |ref|); parameter
and input gradients within 2e-3 * max|grad| (FP32 accumulation over the
batch); round trips within 1e-4 * (1 + |x|).
"""

import ctypes as C

import pytest
import torch

import flowmc_amd._native as N
from flowmc_amd import random as fr
from flowmc_amd.resource.model.nf_model.realNVP import AffineCoupling, RealNVP
from flowmc_amd.resource.optimizer import adam
from oracle import realnvp as R

pytestmark = pytest.mark.gpu

CASES = [(2, 4, 32), (3, 2, 4), (5, 3, 16), (64, 6, 256), (7, 5, 300), (1, 2, 8)]


def make(d, L, h, seed=0):
    # output weights U(+-1.5/sqrt(h)) (reference bound 1/sqrt(h)): well away from the identity,
    # but no layer amplifies rounding by much more than the next
    p = R.init_params(d, L, h, seed=seed, scale=0.5, out_scale=1.5 / h ** 0.5)
    g = torch.Generator().manual_seed(seed + 1)
    p["data_mean"] = torch.randn(d, generator=g, dtype=torch.float64) * 0.5
    p["data_cov"] = torch.diag(torch.rand(d, generator=g, dtype=torch.float64) + 0.5) + 0.01  # off-diagonals ignored
    p["base_mean"] = torch.full((d,), 0.1, dtype=torch.float64)
    p["base_cov"] = torch.eye(d, dtype=torch.float64) * 0.8
    theta = R.to_flat(p).float()
    p = R.from_flat(theta.double(), d, L, h)
    return RealNVP(d, L, h, fr.PRNGKey(0), _theta=theta), p


def close(a, b, tol):
    a = a.detach().cpu().double()
    err = ((a - b).abs() / (1 + b.abs())).max()
    if not bool(err <= tol):
        print(f"max scaled error {float(err):.3e} (tol {tol:.1e}), max |ref| {float(b.abs().max()):.3e}")
    return bool(err <= tol)


@pytest.mark.parametrize("case", CASES)
def test_passes_match_oracle(gpu, case):
    d, L, h = case
    m, p = make(d, L, h)
    g = torch.Generator().manual_seed(2)
    x = torch.randn(333, d, generator=g, dtype=torch.float64) * 2
    y, ld = m.forward(x.float().cuda())
    ry, rld = R.forward(p, x.float().double())
    assert close(y, ry, 2e-4) and close(ld, rld, 2e-4)
    xi, ldi = m.inverse(x.float().cuda())
    rxi, rldi = R.inverse(p, x.float().double())
    assert close(xi, rxi, 2e-4) and close(ldi, rldi, 2e-4)
    assert close(m.log_prob(x.float().cuda()), R.log_prob(p, x.float().double()), 2e-4)
    z = torch.randn(250, d, generator=g)
    assert close(m.sample_from_noise(z.cuda()), R.sample_from_noise(p, z.double()), 2e-4)
    # reference test_nf.py:24-52: inverse recovers x, log-dets cancel
    xr, ldr = m.inverse(y)
    assert close(xr, x.float().double(), 1e-4) and close(ldr, -ld.double().cpu(), 1e-4)


@pytest.mark.parametrize("case", CASES[:4])
def test_adjoint_matches_oracle_autograd(gpu, case):
    d, L, h = case
    m, p = make(d, L, h, seed=3)
    g = torch.Generator().manual_seed(4)
    B = 300
    x = (torch.randn(B, d, generator=g) * 1.5).float()
    w = torch.randn(B, generator=g).float()
    gt, gx = m.grad_log_prob_sum(x.cuda(), w.cuda())
    theta = R.to_flat(p).clone().requires_grad_(True)
    xr = x.double().clone().requires_grad_(True)
    obj = (w.double() * R.log_prob(R.from_flat(theta, d, L, h), xr)).sum()
    rg, rgx = torch.autograd.grad(obj, [theta, xr])
    n = L * 2 * (2 * h * d + h + d)
    gt = gt.cpu().double()
    assert float(gt[n:].abs().max()) == 0.0  # statistics get no gradient
    assert float((gt[:n] - rg[:n]).abs().max()) <= 2e-3 * float(rg[:n].abs().max())
    assert float((gx.cpu().double() - rgx).abs().max()) <= 2e-3 * float(rgx.abs().max())


def test_log_prob_is_differentiable_in_x(gpu):
    m, p = make(4, 3, 16, seed=5)
    x = torch.randn(50, 4, device="cuda", requires_grad=True)
    (gx,) = torch.autograd.grad(m.log_prob(x).sum(), x)
    xr = x.detach().cpu().double().requires_grad_(True)
    (rgx,) = torch.autograd.grad(R.log_prob(p, xr).sum(), xr)
    assert float((gx.cpu().double() - rgx).abs().max()) <= 2e-3 * float(rgx.abs().max())


def test_training_reduces_the_loss_and_samples(gpu):
    # test/integration/test_normalizingFlow.py:12-29 shapes (RealNVP(2, 4, 32), 100
    # points, adam(1e-3, 0.9)), trained long enough to see the loss fall
    key1, rng, init_rng = fr.split(fr.PRNGKey(0), 3)
    g = torch.Generator(device="cuda").manual_seed(0)
    data = torch.randn(2000, 2, device="cuda", generator=g) * torch.tensor([2.0, 0.5], device="cuda") + 1.0
    data[:, 1] += 0.3 * data[:, 0] ** 2
    model = RealNVP(2, 4, 32, rng)
    opt = adam(1e-3, 0.9)
    state = opt.init(model)
    rng, best, state, losses = model.train(init_rng, data, opt, state, 40, 100, verbose=False)
    assert losses.shape == (40,) and float(losses[-1]) < float(losses[0]) - 0.2
    smp = best.sample(fr.PRNGKey(124098), 10000)
    assert smp.shape == (10000, 2) and torch.isfinite(smp).all()
    assert abs(float(smp[:, 0].mean()) - 1.0) < 0.3


def test_affine_coupling_forward_and_inverse(gpu):
    # test_nf.py:8-21
    mask = torch.where(torch.arange(2) % 2 == 0, 1.0, 0.0)
    layer = AffineCoupling(2, 4, mask, fr.PRNGKey(0), 0.5, scale=0.5)
    x = torch.tensor([[1.0, 2.0], [3.0, 4.0]], device="cuda")
    y, ldf = layer.forward(x)
    xr, ldi = layer.inverse(y)
    assert torch.allclose(x, torch.round(xr * 1e5) / 1e5)
    assert torch.allclose(ldf, -ldi)
    # against the oracle: MaskedCouplingLayer(MLPAffine) with conditioning = 1 - mask, dt
    th = layer.theta.cpu().double()
    lp = R.from_flat(th, 2, 1, 4)["layers"][0]
    ry, rld = R.affine_coupling(lp, x.cpu().double(), 1 - mask.double(), dt=0.5)
    assert close(y, ry, 2e-4) and close(ldf, rld, 2e-4)
    assert torch.equal(y[:, 1], x[:, 1])  # mask 0: unchanged


def test_invalid_arguments_are_refused(gpu):
    m, _ = make(3, 2, 4)
    s = m._struct()
    with pytest.raises(N.FlowMCNativeError, match="base log-prob"):
        N.call("fmc_rnvp_pass", N.stream_ptr(), C.byref(s), 1, 2, m.theta.data_ptr(), 1, 0, 0, 0, None, None, None,
               None, None)
    big = N.fmc_rnvp(64, 2, 4096, 1.0, m.theta.data_ptr(), m.packed.data_ptr(), m.layout.data_ptr())
    with pytest.raises(N.FlowMCNativeError, match="LDS"):
        N.call("fmc_rnvp_sizes", C.byref(big), None, None, None, None)
