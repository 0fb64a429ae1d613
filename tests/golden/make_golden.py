"""Writes the golden vectors of tests/golden/*.npz from the FP64 oracle.

Run from the repository root: python tests/golden/make_golden.py
The fixtures are inputs and FP64 outputs of the oracle restatement (pinned
by tests/test_oracle_*.py); the GPU tests compare the kernels with them.
"""

from __future__ import annotations

import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import chains as OC  # noqa: E402
from oracle import flow as F  # noqa: E402
from oracle import optim as O  # noqa: E402
from oracle import targets as OT  # noqa: E402


def f32(t):
    return t.to(torch.float32).to(torch.float64)


def spline():
    K = 8
    g = torch.Generator().manual_seed(100)
    raw = f32(torch.randn(64, 3 * K + 1, generator=g, dtype=torch.float64))
    x = f32(torch.linspace(-12, 12, 64, dtype=torch.float64))
    xp, yp, s = F.spline_params(raw, K)
    y, ld = F.spline_fwd(x, xp, yp, s)
    xi, ldi = F.spline_inv(x, xp, yp, s)
    return dict(K=K, raw=raw.numpy(), x=x.numpy(), y=y.numpy(), ld=ld.numpy(), x_inv=xi.numpy(), ld_inv=ldi.numpy())


def flow():
    d, L, K, h1, h2 = 4, 3, 6, 32, 32
    p = F.init_params(d, L, K, h1, h2, seed=101, scale=0.2)
    theta = f32(F.to_flat(p))
    p = F.from_flat(theta, d, L, K, h1, h2)
    g = torch.Generator().manual_seed(102)
    x = f32(torch.randn(32, d, generator=g, dtype=torch.float64) * 2)
    z = f32(torch.randn(32, d, generator=g, dtype=torch.float64))
    lp = F.log_prob(p, x, K)
    y, ld = F.forward(p, x, K)
    xi, ldi = F.inverse(p, x, K)
    smp = F.sample_from_noise(p, z, K)
    th = theta.clone().requires_grad_(True)
    (gt,) = torch.autograd.grad(F.loss(F.from_flat(th, d, L, K, h1, h2), x, K), th)
    return dict(dims=np.array([d, L, K, h1, h2]), theta=theta.numpy(), x=x.numpy(), z=z.numpy(), log_prob=lp.numpy(),
                y=y.numpy(), ld=ld.numpy(), x_inv=xi.numpy(), ld_inv=ldi.numpy(), sample=smp.numpy(),
                loss_grad=gt.numpy())


def mala():
    d, C, S, dt = 5, 48, 100, 0.3
    g = torch.Generator().manual_seed(103)
    A = torch.randn(d, d, generator=g, dtype=torch.float64)
    P = f32(A @ A.T / d + torch.eye(d, dtype=torch.float64))
    mean = f32(torch.randn(d, generator=g, dtype=torch.float64) * 0.3)
    x0 = f32(torch.randn(C, d, generator=g, dtype=torch.float64))
    noise = f32(torch.randn(S, C, d, generator=g, dtype=torch.float64))
    lu = f32(torch.log(torch.rand(S, C, generator=g, dtype=torch.float64)))
    fn = lambda z: OT.gauss_dense(z, mean, P)
    pos, lps, acc, last, ratios = OC.take_serial_steps("mala", x0, fn, noise, lu, 1, step_size=dt)
    return dict(step_size=dt, mean=mean.numpy(), precision=P.numpy(), x0=x0.numpy(), noise=noise.numpy(),
                log_u=lu.numpy(), positions=pos.numpy(), log_probs=lps.numpy(), accepts=acc.numpy(),
                ratios=ratios.numpy())


def hmc():
    d, C, S, eps, Lf = 8, 32, 30, 0.1, 6
    g = torch.Generator().manual_seed(104)
    x0 = f32(torch.randn(C, d, generator=g, dtype=torch.float64) * 0.5)
    noise = f32(torch.randn(S, C, d, generator=g, dtype=torch.float64))
    lu = f32(torch.log(torch.rand(S, C, generator=g, dtype=torch.float64)))
    metric = f32(0.5 + torch.rand(d, generator=g, dtype=torch.float64))
    pos, lps, acc, last, ratios = OC.take_serial_steps("hmc", x0, OT.funnel, noise, lu, 1, step_size=eps,
                                                       n_leapfrog=Lf, metric=metric)
    return dict(step_size=eps, n_leapfrog=Lf, metric=metric.numpy(), x0=x0.numpy(), noise=noise.numpy(),
                log_u=lu.numpy(), positions=pos.numpy(), log_probs=lps.numpy(), accepts=acc.numpy(),
                ratios=ratios.numpy())


def nf_mh():
    C, S, d = 40, 30, 3
    g = torch.Generator().manual_seed(105)
    x0 = f32(torch.randn(C, d, generator=g, dtype=torch.float64))
    lp0, q0 = f32(torch.randn(C, generator=g, dtype=torch.float64)), f32(torch.randn(C, generator=g, dtype=torch.float64))
    prop = f32(torch.randn(C, S, d, generator=g, dtype=torch.float64))
    lpp = f32(torch.randn(C, S, generator=g, dtype=torch.float64))
    qp = f32(torch.randn(C, S, generator=g, dtype=torch.float64))
    lu = f32(torch.log(torch.rand(C, S, generator=g, dtype=torch.float64)))
    pos, lps, acc = OC.nf_mh_scan(x0, lp0, q0, prop, lpp, qp, lu)
    return dict(x0=x0.numpy(), lp0=lp0.numpy(), q0=q0.numpy(), prop=prop.numpy(), lp_prop=lpp.numpy(),
                q_prop=qp.numpy(), log_u=lu.numpy(), positions=pos.numpy(), log_probs=lps.numpy(),
                accepts=acc.numpy())


def adamw():
    n = 257
    g = torch.Generator().manual_seed(106)
    theta = f32(torch.randn(n, generator=g, dtype=torch.float64))
    grads = f32(torch.randn(3, n, generator=g, dtype=torch.float64) * torch.tensor([[0.2], [0.01], [0.5]]))
    th, m, v = theta.clone(), torch.zeros(n, dtype=torch.float64), torch.zeros(n, dtype=torch.float64)
    out = []
    for t in range(3):
        th, m, v = O.adamw_clip_step(th, grads[t], m, v, t + 1, 1e-2, max_norm=1.0)
        out.append(th.clone())
    return dict(theta=theta.numpy(), grads=grads.numpy(), thetas=torch.stack(out).numpy(), lr=1e-2)


BUILDERS = {"spline": spline, "flow": flow, "mala": mala, "hmc": hmc, "nf_mh": nf_mh, "adamw": adamw}


def build_all():
    return {k: f() for k, f in BUILDERS.items()}


if __name__ == "__main__":
    for name, arrays in build_all().items():
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrays)
        print("wrote", name)
