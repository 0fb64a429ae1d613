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
