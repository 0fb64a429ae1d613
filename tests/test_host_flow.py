"""Host-side flow layout: the C++ layout (fmc_flow_sizes) and the Python /
oracle flat-parameter layouts agree; invalid flow descriptions are rejected."""

import ctypes as C

import pytest
import torch

import flowmc_amd._native as N
from oracle import flow as F


def flow(d, L, K, hidden, prec=1, theta=1, packed=None, pbias=None):
    return N.make_flow(d, L, K, hidden, prec, -10.0, 10.0, 1e-4, 1e-4, theta, packed, pbias)


def sizes(d, L, K, h1, h2=None, prec=1):
    f = flow(d, L, K, [h1, h2] if h2 is not None else h1, prec)
    out = [C.c_int64() for _ in range(4)] + [C.c_int32(), C.c_int32()]
    N.call("fmc_flow_sizes", C.byref(f), *[C.byref(o) for o in out])
    return [o.value for o in out]


@pytest.mark.parametrize("case", [(5, 4, 8, [32, 32]), (32, 8, 10, [128, 128]), (64, 10, 8, [256, 256]),
                                  (7, 3, 4, [16, 48]), (6, 2, 8, [32]), (9, 3, 8, [64, 32, 48]),
                                  (8, 2, 12, [32, 32, 32, 32])])
def test_theta_layout_matches_oracle(case):
    d, L, K, hidden = case
    th, pk, pb, rows, tf, tb = sizes(d, L, K, hidden)
    p = F.init_params(d, L, K, hidden)
    assert th == F.to_flat(p).numel()
    assert pk > 0 and pb > 0 and rows > 0 and tf in (32, 64) and tb in (32, 64)
    # store rows: conditioner input, every hidden layer and its gradient, spline-parameter gradients
    P, dt = 3 * K + 1, d - d // 2
    n3 = 64 * ((dt + 1) // 2) if K <= 10 else None
    if n3 is not None:
        hp = [(h + 31) // 32 * 32 for h in hidden]
        assert rows == 32 * ((d + 1) // 2 + 31 >= 32) * (((d + 1) // 2 + 31) // 32) + 2 * sum(hp) + n3


def test_packed_sizes_with_one_and_three_hidden_layers():
    # bf16, d=64, K=8: per layer wh(i) / wht(i) images plus the register-order output images
    _, pk1, pb1, rows1, _, _ = sizes(64, 2, 8, [256])
    assert pk1 == 2 * (2 * 256 * 32 + 2 * 1024 * 256)
    assert pb1 == 2 * (256 + 1024)
    _, pk3, pb3, rows3, _, _ = sizes(64, 2, 8, [256, 128, 64])
    per = 2 * 256 * 32 + 2 * 128 * 256 + 2 * 64 * 128 + 2 * 1024 * 64
    assert pk3 == 2 * per
    assert pb3 == 2 * (256 + 128 + 64 + 1024)
    assert rows3 == 32 + 2 * (256 + 128 + 64) + 1024


def test_packed_sizes_scale_with_transformed_dims():
    # compact conditioner: only conditioning columns of W1 and transformed rows of W3 are packed.
    # d=64, K=8 (P=25), hidden 256, bf16 (K granularity 16), per layer:
    #   w1 / w1t: 256 x 32 each; w2 / w2t: 256 x 256 each;
    #   w3r / w3rt: 16 dim pairs x 64 rows x 256 each (register-spline order, 32 slots per dim)
    th, pk, pb, rows, _, _ = sizes(64, 2, 8, 256, 256)
    per_layer = 2 * 256 * 32 + 2 * 256 * 256 + 2 * 1024 * 256
    assert pk == 2 * per_layer
    assert pb == 2 * (256 + 256 + 1024)
    assert rows == 32 + 256 + 256 + 256 + 256 + 1024  # c, h1, h2, gp1, gp2, gp
    # K > 10: chunked images w3 / w3t only (chunks of <= 3 dims, 10 x 128 + 96 rows)
    _, pk11, pb11, _, _, _ = sizes(64, 2, 11, 256, 256)
    P = 34
    G = 128 // P
    n3p = (32 // G) * 128 + (((32 % G) * P + 31) // 32) * 32 if 32 % G else (32 // G) * 128
    assert pk11 == 2 * (2 * 256 * 32 + 2 * 256 * 256 + 2 * n3p * 256)
    full = 2 * (2 * (256 * 64 + 256 * 256 + 64 * 25 * 256))  # forward + transposed images at full width
    assert pk < full


def test_invalid_flow_rejected():
    f = flow(1, 2, 8, [32, 32])
    z = C.c_int64()
    with pytest.raises(N.FlowMCNativeError, match="n_dim"):
        N.call("fmc_flow_sizes", C.byref(f), C.byref(z), None, None, None, None, None)
    f = flow(4, 2, 8, [32, 32], prec=7)
    with pytest.raises(N.FlowMCNativeError, match="precision"):
        N.call("fmc_flow_sizes", C.byref(f), C.byref(z), None, None, None, None, None)
    f = flow(4, 2, 8, [32, 32])
    with pytest.raises(N.FlowMCNativeError, match="packed buffers"):
        N.call("fmc_flow_pass", None, C.byref(f), 0, 0, 1, 4, 0, 0, 0, None, None, None, None, None)
    # the base log-prob of a reference-order inverse pass is refused (its output
    # does not have the forward map's density); checked before any launch
    f = flow(4, 2, 8, [32, 32], packed=16, pbias=24, theta=8)
    with pytest.raises(N.FlowMCNativeError, match="base log-prob"):
        N.call("fmc_flow_pass", None, C.byref(f), 1, 2 | 4, 0, 4, 0, 0, 0, None, None, None, None, None)


def test_hidden_layer_count_is_checked():
    f = flow(4, 2, 8, [32])
    f.n_hidden = 5
    with pytest.raises(N.FlowMCNativeError, match="1 to 4 hidden layers"):
        N.call("fmc_flow_sizes", C.byref(f), None, None, None, None, None, None)
    f.n_hidden = 0
    with pytest.raises(N.FlowMCNativeError, match="1 to 4 hidden layers"):
        N.call("fmc_flow_sizes", C.byref(f), None, None, None, None, None, None)
    with pytest.raises(N.FlowMCNativeError, match="hidden size 2"):
        N.call("fmc_flow_sizes", C.byref(flow(4, 2, 8, [32, 32, 600])), None, None, None, None, None, None)
    with pytest.raises(ValueError):
        N.make_flow(4, 2, 8, [], 1)


def test_gaussian_view_log_prob_uses_the_full_covariance():
    # Gaussian.log_prob = multivariate_normal.logpdf (common.py:285-286), dense covariance
    from flowmc_amd.resource.model.common import Gaussian

    g = torch.Generator().manual_seed(3)
    A = torch.randn(5, 5, generator=g, dtype=torch.float64)
    cov = A @ A.T / 5 + 0.3 * torch.eye(5, dtype=torch.float64)
    mean = torch.randn(5, generator=g, dtype=torch.float64)
    x = torch.randn(20, 5, generator=g, dtype=torch.float64)
    want = torch.distributions.MultivariateNormal(mean, cov).log_prob(x)
    got = Gaussian(mean, cov).log_prob(x)
    assert torch.allclose(got.double(), want, atol=1e-4, rtol=1e-5)


def test_realnvp_refuses_a_learnable_base():
    # its kernels need a diagonal base, which training would leave; the check
    # comes before any device work
    from flowmc_amd import random as fr
    from flowmc_amd.resource.model.common import Gaussian
    from flowmc_amd.resource.model.nf_model.realNVP import RealNVP

    with pytest.raises(NotImplementedError, match="learnable"):
        RealNVP(4, 2, 16, fr.PRNGKey(0), base_dist=Gaussian(torch.zeros(4), torch.eye(4), learnable=True))


@pytest.mark.parametrize("n", [2, 255, 3001, 70000])
def test_data_stats_match_numpy(n):
    """train()'s data mean and covariance (jnp.mean, jnp.cov with ddof 1,
    nf_model/base.py:187-188): chunked batched products plus the ragged tail
    against numpy's FP64 np.cov."""
    import numpy as np

    from flowmc_amd.resource.model.nf_model.base import data_stats

    g = torch.Generator().manual_seed(n)
    x = torch.randn(n, 5, generator=g) * torch.tensor([1.0, 2.0, 3.0, 0.5, 4.0]) + 3.0
    x[:, 1] += 0.7 * x[:, 0]  # correlated columns: off-diagonal terms
    mean, cov = data_stats(x)
    xd = x.double().numpy()
    np.testing.assert_allclose(mean.numpy(), xd.mean(0), rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(cov.numpy(), np.cov(xd.T), rtol=1e-4, atol=1e-4)
