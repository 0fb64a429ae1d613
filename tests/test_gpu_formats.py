"""Flow files: save_model / load_model round trips of MaskedCouplingRQSpline
and RealNVP (flat FP32 parameters + configuration in one .npz; the
reference's Equinox .eqx leaf files are not read)."""

import pytest
import torch

from flowmc_amd import random as fr
from flowmc_amd.resource.model.nf_model.realNVP import RealNVP
from flowmc_amd.resource.model.nf_model.rqSpline import MaskedCouplingRQSpline

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["rqspline", "realnvp"])
def test_flow_save_load_round_trip(gpu, tmp_path, kind):
    def make(L, seed):
        if kind == "rqspline":
            return MaskedCouplingRQSpline(5, L, [16, 16], 8, fr.PRNGKey(seed))
        return RealNVP(5, L, 16, fr.PRNGKey(seed))

    m = make(3, 0)
    g = torch.Generator(device="cuda").manual_seed(0)
    m.theta.add_(torch.randn(m.theta.shape, device="cuda", generator=g) * 0.05)
    m.mark_updated()
    path = str(tmp_path) + "/flow"
    m.save_resource(path)
    m2 = make(3, 1).load_resource(path)
    x = torch.randn(200, 5, device="cuda", generator=g)
    assert torch.equal(m.log_prob(x), m2.log_prob(x))
    assert torch.equal(m.sample(fr.PRNGKey(3), 50), m2.sample(fr.PRNGKey(3), 50))
    with pytest.raises(ValueError):
        make(2, 0).load_resource(path)
