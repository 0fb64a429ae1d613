"""Pins the oracle's MALA / HMC / NF-MH restatements to properties the
mathematics fixes: detailed balance of MALA (densities from scipy), leapfrog
reversibility and energy behaviour, gradient-call counts, acceptance -> 1."""

import math

import numpy as np
import pytest
import scipy.stats as ss
import torch

from oracle import chains, targets as T



@pytest.fixture(autouse=True, scope="module")
def _float64_default():
    old = torch.get_default_dtype()
    torch.set_default_dtype(torch.float64)
    yield
    torch.set_default_dtype(old)


def test_mala_detailed_balance_gaussian():
    # pi = N(mu, Sigma); MALA transition density from x to y is
    # q(y|x) * min(1, e^{r(x,y)}); detailed balance needs
    # pi(x) q(y|x) a(x,y) == pi(y) q(x|y) a(y,x)
    rng = np.random.default_rng(0)
    d, dt = 3, 0.7
    A = rng.normal(size=(d, d))
    cov = A @ A.T + np.eye(d)
    P = np.linalg.inv(cov)
    mu = rng.normal(size=d)
    fn = lambda z: T.gauss_dense(z, torch.tensor(mu), torch.tensor(P))
    for _ in range(20):
        x = torch.tensor(rng.normal(size=(1, d)))
        xi = torch.tensor(rng.normal(size=(1, d)))
        _, _, _, r_xy = chains.mala_step(x, fn, dt, xi, torch.tensor([-1e300]))
        y = x + dt * dt / 2 * torch.tensor((-(x.numpy() - mu) @ P)) + dt * xi
        xi_back = (x - y - dt * dt / 2 * torch.tensor(-(y.numpy() - mu) @ P)) / dt
        _, _, _, r_yx = chains.mala_step(y, fn, dt, xi_back, torch.tensor([-1e300]))
        grad = lambda z: -(z - mu) @ P
        q_yx = ss.multivariate_normal(x.numpy()[0] + dt * dt / 2 * grad(x.numpy()[0]), dt * dt * np.eye(d)).logpdf(
            y.numpy()[0])
        q_xy = ss.multivariate_normal(y.numpy()[0] + dt * dt / 2 * grad(y.numpy()[0]), dt * dt * np.eye(d)).logpdf(
            x.numpy()[0])
        pi = ss.multivariate_normal(mu, cov)
        lhs = pi.logpdf(x.numpy()[0]) + q_yx + min(0.0, float(r_xy))
        rhs = pi.logpdf(y.numpy()[0]) + q_xy + min(0.0, float(r_yx))
        assert abs(lhs - rhs) < 1e-9
        assert abs(float(r_xy) + float(r_yx)) < 1e-9


def test_mala_acceptance_goes_to_one():
    x = torch.randn(100, 2)
    xi = torch.randn(100, 2)
    lu = torch.log(torch.rand(100))
    fn = lambda z: -0.5 * (z * z).sum(-1)
    _, _, acc, _ = chains.mala_step(x, fn, 1e-5, xi, lu)
    assert bool(acc.all())


def _leapfrog_only(x, p, fn, eps, L):
    # the coefficient table of HMC.py:47-50 without the MH step
    coefs = torch.ones(L + 2, 2)
    coefs[0] = torch.tensor([0.0, 0.5])
    coefs[-1] = torch.tensor([1.0, 0.5])
    for i in range(L + 2):
        x = x + eps * coefs[i, 0] * p
        _, g = T.value_and_grad(fn, x)
        p = p + eps * coefs[i, 1] * g
    return x, p


def test_hmc_leapfrog_reversible_and_accepts_small_steps():
    fn = lambda z: -0.5 * (z * z).sum(-1)
    x0 = torch.randn(5, 2)
    p0 = torch.randn(5, 2)
    x1, p1 = _leapfrog_only(x0, p0, fn, 1.0, 5)
    x2, p2 = _leapfrog_only(x1, -p1, fn, 1.0, 5)
    assert torch.allclose(x2, x0, atol=1e-12) and torch.allclose(-p2, p0, atol=1e-12)
    lp = fn(x0)
    _, _, acc, _ = chains.hmc_step(x0, lp, fn, 1e-7, 5, torch.ones(2), torch.randn(5, 2), torch.log(torch.rand(5)))
    assert bool(acc.all())


def test_hmc_energy_error_is_second_order():
    # harmonic oscillator: leapfrog H error scales as eps^2 (halving eps -> ~1/4)
    fn = lambda z: -0.5 * (z * z).sum(-1)
    x0 = torch.tensor([[1.0, -0.5]])
    xi = torch.tensor([[0.3, 0.8]])
    errs = []
    for eps in (0.1, 0.05):
        L = int(round(1.0 / eps))
        _, _, _, log_acc = chains.hmc_step(x0, fn(x0), fn, eps, L, torch.ones(2), xi, torch.tensor([0.0]))
        errs.append(abs(float(log_acc)))
    assert 3.0 < errs[0] / errs[1] < 5.0


def test_hmc_gradient_call_count():
    calls = []

    def fn(z):
        calls.append(1)
        return -0.5 * (z * z).sum(-1)

    x = torch.randn(1, 3)
    lp = fn(x)
    calls.clear()
    chains.hmc_step(x, lp, fn, 0.1, 10, torch.ones(3), torch.randn(1, 3), torch.tensor([0.0]))
    # n_leapfrog + 2 gradient evaluations (HMC.py:91-95) + 1 potential (HMC.py:141)
    assert len(calls) == 10 + 2 + 1


def test_nf_mh_accepts_everything_when_target_is_the_flow():
    C, S, d = 4, 7, 3
    prop = torch.randn(C, S, d)
    q = torch.randn(C, S)
    x0 = torch.randn(C, d)
    q0 = torch.randn(C)
    pos, lps, acc = chains.nf_mh_scan(x0, q0, q0, prop, q, q, torch.log(torch.rand(C, S)))
    assert bool(acc.all())
    assert torch.equal(pos, prop)


def test_take_serial_steps_thinning_rows():
    fn = lambda z: -0.5 * (z * z).sum(-1)
    S, C, d = 6, 3, 2
    noise = torch.randn(S, C, d)
    lu = torch.log(torch.rand(S, C))
    pos, lps, accs, last, _ = chains.take_serial_steps("mala", torch.zeros(C, d), fn, noise, lu, thinning=4,
                                                       step_size=0.5)
    full, _, _, _, _ = chains.take_serial_steps("mala", torch.zeros(C, d), fn, noise, lu, thinning=1, step_size=0.5)
    assert pos.shape == (C, 2, d)
    assert torch.equal(pos[:, 1], full[:, 4]) and torch.equal(last, full[:, 4])


def test_dynamic_update_slice_clamps():
    buf = torch.zeros(2, 5)
    out = chains.dynamic_update_slice(buf, torch.ones(2, 3), 4, 1)
    assert torch.equal(out[:, 2:], torch.ones(2, 3)) and float(out[:, :2].abs().sum()) == 0


def test_grw_detailed_balance_gaussian():
    # symmetric proposal: pi(x) q(y|x) min(1, e^r) == pi(y) q(x|y) min(1, e^-r),
    # with pi and q from scipy (independent of the oracle's target code)
    rng = np.random.default_rng(1)
    d, step = 3, 0.8
    A = rng.normal(size=(d, d))
    cov = A @ A.T + np.eye(d)
    P = np.linalg.inv(cov)
    mu = rng.normal(size=d)
    fn = lambda z: T.gauss_dense(z, torch.tensor(mu), torch.tensor(P))
    pi = ss.multivariate_normal(mu, cov)
    for _ in range(20):
        x = torch.tensor(rng.normal(size=(1, d)))
        xi = torch.tensor(rng.normal(size=(1, d)))
        lp_x = torch.tensor([float(T.value_and_grad(fn, x)[0][0])])
        _, _, _, r = chains.grw_step(x, lp_x, fn, step, xi, torch.tensor([-1e300]))
        y = (x + step * xi).numpy()[0]
        q = ss.multivariate_normal(np.zeros(d), step * step * np.eye(d))
        lhs = pi.pdf(x.numpy()[0]) * q.pdf(y - x.numpy()[0]) * min(1.0, math.exp(float(r[0])))
        rhs = pi.pdf(y) * q.pdf(x.numpy()[0] - y) * min(1.0, math.exp(-float(r[0])))
        assert lhs == pytest.approx(rhs, rel=1e-9)
        # the ratio is the log-density difference (the proposal is symmetric)
        assert float(r[0]) == pytest.approx(pi.logpdf(y) - pi.logpdf(x.numpy()[0]), rel=1e-9, abs=1e-9)


def test_grw_uses_the_passed_log_prob_and_accepts_tiny_steps():
    g = torch.Generator().manual_seed(2)
    fn = lambda z: -0.5 * (z * z).sum(-1)
    x = torch.randn(200, 2, generator=g)
    lp = fn(x)
    xi = torch.randn(200, 2, generator=g)
    lu = torch.log(torch.rand(200, generator=g))
    _, _, acc, _ = chains.grw_step(x, lp, fn, 1e-5, xi, lu)
    assert bool(acc.all())
    # a log_prob far above the truth makes every proposal look worse: nothing accepted
    _, lp2, acc2, _ = chains.grw_step(x, lp + 1e3, fn, 1e-5, xi, lu)
    assert not bool(acc2.any()) and torch.equal(lp2, lp + 1e3)


def test_grw_moments_standard_normal():
    # reference test_kernels.py:352-392 (mean 0, variance 1 within 3e-2), many chains
    g = torch.Generator().manual_seed(3)
    C, d, S = 4000, 2, 400
    fn = lambda z: -0.5 * (z * z).sum(-1)
    x = torch.randn(C, d, generator=g)
    noise = torch.randn(S, C, d, generator=g)
    lu = torch.log(torch.rand(S, C, generator=g))
    pos, _, _, _, _ = chains.take_serial_steps("grw", x, fn, noise, lu, step_size=1.0)
    tail = pos[:, 100:]
    assert abs(float(tail.mean())) < 3e-2 and abs(float(tail.var()) - 1.0) < 3e-2


# ---------------------------------------------------------------------------
# Non-unit condition matrices (HMC.py:118-136). Two independent pins:
#   (1) convention: on a flat target no kick moves p, so the L+1 drifts
#       displace x by eps (L+1) * dK/dp, whose covariance is fixed in closed
#       form by the momentum law p ~ N(0, M^-1) (the momentum draw
#       xi @ chol(inv(M)).T) and the broadcast kinetic 1/2 sum(p^2 * M);
#   (2) consistency: for a diagonal M, one HMC step started from exact draws
#       of a Gaussian target keeps its mean and variance (the joint
#       exp(-U - K) has the momentum marginal the step draws from).
# Each is checked to reject the planted mistake p = xi * sqrt(m).


def _hmc_displacement_cov(step, metric, C=200_000, d=4, eps=0.1, L=4, seed=0):
    g = torch.Generator().manual_seed(seed)
    x = torch.zeros(C, d)
    xi = torch.randn(C, d, generator=g)
    flat = lambda z: 0.0 * z.sum(-1)
    xn, _, acc, log_acc = step(x, flat(x), flat, eps, L, metric, xi, torch.full((C,), -1e300))
    assert bool(acc.all()) and float(log_acc.abs().max()) < 1e-12  # p is unchanged: H0 == H1
    disp = (xn - x) / (eps * (L + 1))
    return torch.cov(disp.T)


def _swapped_momentum_hmc_step(x, lp, fn, eps, n_leapfrog, metric, xi, log_u):
    # the planted mistake, on the CPU: momentum xi * sqrt(m) instead of xi / sqrt(m)
    return chains.hmc_step(x, lp, fn, eps, n_leapfrog, metric, xi * metric, log_u)


def test_hmc_diagonal_metric_displacement_covariance():
    m = torch.tensor([0.3, 2.5, 1.0, 4.0])
    want = torch.diag(m)  # dK/dp = m p with p ~ N(0, diag(1/m)): Cov = M M^-1 M = M
    got = _hmc_displacement_cov(chains.hmc_step, m)
    assert torch.allclose(got, want, atol=0.02 * float(m.max()), rtol=0.02)
    bad = _hmc_displacement_cov(_swapped_momentum_hmc_step, m)
    assert not torch.allclose(bad, want, atol=0.02 * float(m.max()), rtol=0.02)


def test_hmc_dense_metric_displacement_covariance():
    # dense SPD M: p ~ N(0, M^-1) (its Cholesky factor from numpy, not torch);
    # the broadcast kinetic's p-gradient is p * colsum(M)
    rng = np.random.default_rng(4)
    A = rng.normal(size=(4, 4))
    M = A @ A.T / 4 + np.diag([0.5, 1.0, 2.0, 3.0])
    c = M.sum(0)
    want = torch.tensor(np.diag(c) @ np.linalg.inv(M) @ np.diag(c))
    Lnp = np.linalg.cholesky(np.linalg.inv(M))
    assert np.allclose(Lnp @ Lnp.T, np.linalg.inv(M))
    got = _hmc_displacement_cov(chains.hmc_step, torch.tensor(M))
    tol = 0.02 * float(want.abs().max())
    assert torch.allclose(got, want, atol=tol, rtol=0.02)
    # a diagonal matrix given densely is the diagonal path, bit for bit
    g = torch.Generator().manual_seed(6)
    fn = lambda z: T.gauss_diag(z, torch.zeros(4), torch.tensor([1.0, 2.0, 0.5, 1.5]))
    x = torch.randn(16, 4, generator=g)
    xi = torch.randn(16, 4, generator=g)
    lu = torch.log(torch.rand(16, generator=g))
    md = torch.tensor([0.7, 1.3, 2.0, 0.4])
    a = chains.hmc_step(x, fn(x), fn, 0.2, 3, md, xi, lu)
    b = chains.hmc_step(x, fn(x), fn, 0.2, 3, torch.diag(md), xi, lu)
    assert torch.allclose(a[0], b[0], atol=1e-14) and torch.allclose(a[3], b[3], atol=1e-12)


def _one_step_moments(step, m, C=40_000, seed=1):
    g = torch.Generator().manual_seed(seed)
    mu = torch.tensor([1.0, -2.0, 0.5, 0.0])
    sd = torch.tensor([0.5, 2.0, 1.0, 1.5])
    fn = lambda z: T.gauss_diag(z, mu, 1.0 / sd ** 2)
    x = mu + sd * torch.randn(C, 4, generator=g)  # exact draws from the target
    xi = torch.randn(C, 4, generator=g)
    lu = torch.log(torch.rand(C, generator=g))
    xn, _, acc, _ = step(x, fn(x), fn, 0.35, 4, m, xi, lu)
    z = (xn - mu) / sd
    return z.mean(0), z.var(0), float(acc.float().mean())


def test_hmc_diagonal_metric_keeps_gaussian_moments():
    m = torch.tensor([3.0, 0.4, 2.0, 0.25])
    mean, var, acc = _one_step_moments(chains.hmc_step, m)
    # N(0, 1) after standardisation: 5 sigma of the sampling error at C = 40k
    assert acc > 0.5
    assert float(mean.abs().max()) < 5 / 40_000 ** 0.5 * 1.5
    assert float((var - 1).abs().max()) < 5 * (2 / 40_000) ** 0.5 * 1.5
    _, bvar, _ = _one_step_moments(_swapped_momentum_hmc_step, m)
    assert float((bvar - 1).abs().max()) > 0.1
