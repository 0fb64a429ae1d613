"""Flow files: save_model / load_model round trips of MaskedCouplingRQSpline
and RealNVP (flat FP32 parameters + configuration in one .npz; the
reference's Equinox .eqx leaf files are not read)."""

import pytest
import torch

from flowmc_amd import random as fr
from flowmc_amd.resource.model.nf_model.realNVP import RealNVP
from flowmc_amd.resource.model.nf_model.rqSpline import MaskedCouplingRQSpline

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["rqspline", "rqspline_dense_base", "rqspline_learnable_base", "realnvp"])
def test_flow_save_load_round_trip(gpu, tmp_path, kind):
    # the loading model is a plain one: the file decides the base path
    # (diagonal or full covariance, learnable or not)
    from flowmc_amd.resource.model.common import Gaussian

    def make(L, seed, plain=False):
        if kind.startswith("rqspline"):
            bd = None
            if kind == "rqspline_dense_base" and not plain:
                a = torch.randn(5, 5, generator=torch.Generator().manual_seed(seed))
                bd = Gaussian(torch.zeros(5), a @ a.T / 5 + 0.5 * torch.eye(5))
            elif kind == "rqspline_learnable_base" and not plain:
                bd = Gaussian(torch.zeros(5), torch.eye(5), learnable=True)
            return MaskedCouplingRQSpline(5, L, [16, 16], 8, fr.PRNGKey(seed), base_dist=bd)
        return RealNVP(5, L, 16, fr.PRNGKey(seed))

    m = make(3, 0)
    g = torch.Generator(device="cuda").manual_seed(0)
    m.theta.add_(torch.randn(m.theta.shape, device="cuda", generator=g) * 0.05)
    m.mark_updated()
    path = str(tmp_path) + "/flow"
    m.save_resource(path)
    m2 = make(3, 1, plain=True).load_resource(path)
    if kind.startswith("rqspline"):
        assert (m2._base_dense, m2._base_learnable) == (m._base_dense, m._base_learnable)
    x = torch.randn(200, 5, device="cuda", generator=g)
    assert torch.equal(m.log_prob(x), m2.log_prob(x))
    assert torch.equal(m.sample(fr.PRNGKey(3), 50), m2.sample(fr.PRNGKey(3), 50))
    with pytest.raises(ValueError):
        make(2, 0).load_resource(path)
