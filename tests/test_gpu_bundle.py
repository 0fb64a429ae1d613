"""End-to-end NF-MCMC through Sampler + RQSpline_MALA_Bundle
(test/integration/test_quickstart.py, test/unit/test_bundle.py), with the
assertions the reference leaves out."""

import pytest
import torch

from flowmc_amd import random as fr
from flowmc_amd import targets as TG
from flowmc_amd.resource.logPDF import LogPDF
from flowmc_amd.resource_strategy_bundle.RQSpline_MALA import RQSpline_MALA_Bundle
from flowmc_amd.Sampler import Sampler

pytestmark = pytest.mark.gpu


def log_posterior(x, data: dict):
    return -0.5 * torch.sum((x - data["data"]) ** 2)


@pytest.mark.parametrize("native", [False, True])
def test_quickstart_bundle(gpu, native):
    n_dims, n_chains = 2, 10
    n_local, n_global, n_train, n_prod, n_epochs = 10, 10, 5, 5, 10
    data = {"data": torch.arange(n_dims, dtype=torch.float32, device=gpu)}
    key = fr.PRNGKey(42)
    key, sub = fr.split(key)
    x0 = fr.normal(sub, (n_chains, n_dims))
    key, sub = fr.split(key)
    logpdf = LogPDF(TG.GaussianDiag(torch.arange(n_dims)), n_dims=n_dims) if native else log_posterior
    bundle = RQSpline_MALA_Bundle(sub, n_chains, n_dims, logpdf, n_local, n_global, n_train, n_prod, n_epochs,
                                  rq_spline_hidden_units=[64, 64], rq_spline_n_bins=8, rq_spline_n_layers=3)
    assert repr(bundle) == "RQSpline_MALA Bundle"
    assert len(bundle.strategy_order) == 6 * n_train + 2 + 4 * n_prod
    sampler = Sampler(n_dims, n_chains, key, resource_strategy_bundles=bundle)
    sampler.sample(x0, data)
    r = sampler.resources
    for name in ("positions_training", "log_prob_training", "positions_production", "log_prob_production"):
        assert bool(torch.isfinite(r[name].data).all()), name
    for name in ("local_accs_training", "global_accs_training", "local_accs_production", "global_accs_production"):
        acc = r[name].data
        # local and global steppers share the buffer rows: each writes its own half
        filled = torch.isfinite(acc)
        assert bool(((acc[filled] == 0) | (acc[filled] == 1)).all())
    assert bool(torch.isfinite(r["loss_buffer"].data).all())
    assert r["sampler_state"].data["training"] is False
    pos = r["positions_production"].data
    assert abs(float(pos[..., 0].mean()) - 0.0) < 0.6 and abs(float(pos[..., 1].mean()) - 1.0) < 0.6
    # the global sampler holds the trained model
    assert r["global_sampler"].model is r["model"]


def test_nfmcmc_bundle_at_scale_keeps_the_mixture(gpu):
    """The full-size NF-MCMC configuration through Sampler + bundle: the
    64-D 8-mode mixture, 4096 chains, MALA step 0.05, the 10-layer 8-bin
    [256, 256] BF16 flow trained 3 epochs a round over 100k rows of the last
    200 per chain (batch 10k, lr 2e-4), the exact inverse for sampling, 20
    training + 4 production rounds of 100 local + 100 global steps, every
    chain started from one exact mixture draw. Checked on the production rows:
    every mode keeps 1/8 of the chains within 0.03 (binomial sd 0.005 at 4096
    chains; a mode the sampler loses or floods moves it by 0.125), the
    pooled mean within 0.1 sd and the variance within 10 % of the analytic
    moments in every dimension, the global sampler accepts more than 10 % of
    its proposals (the trained flow proposes into the modes; about 30 % after
    20 rounds) and every buffer is finite."""
    d, C, nl, ng, n_train, n_prod = 64, 4096, 100, 100, 20, 4
    target = TG.gaussian_mixture_64d(seed=0, device=gpu)
    g = torch.Generator(device=gpu).manual_seed(1)
    mu = target.means.to(gpu, torch.float32)
    sg = target.sigmas.to(gpu, torch.float32)
    lab = torch.randint(0, mu.shape[0], (C,), device=gpu, generator=g)
    x0 = (mu[lab] + sg[lab] * torch.randn(C, d, device=gpu, generator=g)).contiguous()
    key = fr.PRNGKey(3)
    key, sub = fr.split(key)
    bundle = RQSpline_MALA_Bundle(sub, C, d, LogPDF(target, n_dims=d), nl, ng, n_train, n_prod, 3,
                                  mala_step_size=0.05, rq_spline_hidden_units=[256, 256], rq_spline_n_bins=8,
                                  rq_spline_n_layers=10, learning_rate=2e-4, batch_size=10000,
                                  n_max_examples=100000, precision="bfloat16", exact_inverse=True,
                                  history_window=200)
    sampler = Sampler(d, C, key, resource_strategy_bundles=bundle)
    sampler.sample(x0, {})
    r = sampler.resources
    pos = r["positions_production"].data
    assert bool(torch.isfinite(pos).all()) and bool(torch.isfinite(r["log_prob_production"].data).all())
    rows = pos.reshape(-1, d)
    occ = torch.bincount(torch.cdist(rows, mu).argmin(1), minlength=mu.shape[0]).double() / rows.shape[0]
    mean_t, var_t = target.mean().to(gpu).double(), target.variance().to(gpu).double()
    m = rows.double().mean(0)
    v = rows.double().var(0)
    mean_err = float(((m - mean_t).abs() / var_t.sqrt()).max())
    var_err = float(((v - var_t) / var_t).abs().max())
    gacc = r["global_accs_production"].data
    gacc = float(gacc[torch.isfinite(gacc)].mean())
    print(f"occupancy {occ.tolist()}, mean err {mean_err:.4f} sd, var err {var_err:.4f}, global acc {gacc:.3f}")
    assert float((occ - 0.125).abs().max()) <= 0.03, occ
    assert mean_err <= 0.1 and var_err <= 0.1, (mean_err, var_err)
    assert gacc > 0.1, gacc
