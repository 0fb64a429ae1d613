"""Host-side flow layout: the C++ layout (fmc_flow_sizes) and the Python /
oracle flat-parameter layouts agree; invalid flow descriptions are rejected."""

import ctypes as C

import pytest
import torch

import flowmc_amd._native as N
from oracle import flow as F


def sizes(d, L, K, h1, h2, prec=1):
    f = N.fmc_flow(d, L, K, h1, h2, prec, -10.0, 10.0, 1e-4, 1e-4, 1, None, None)
    out = [C.c_int64() for _ in range(4)] + [C.c_int32(), C.c_int32()]
    N.call("fmc_flow_sizes", C.byref(f), *[C.byref(o) for o in out])
    return [o.value for o in out]


@pytest.mark.parametrize("case", [(5, 4, 8, 32, 32), (32, 8, 10, 128, 128), (64, 10, 8, 256, 256), (7, 3, 4, 16, 48)])
def test_theta_layout_matches_oracle(case):
    d, L, K, h1, h2 = case
    th, pk, pb, rows, tf, tb = sizes(*case)
    p = F.init_params(d, L, K, h1, h2)
    assert th == F.to_flat(p).numel()
    assert pk > 0 and pb > 0 and rows > 0 and tf in (32, 64) and tb in (32, 64)


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
    f = N.fmc_flow(1, 2, 8, 32, 32, 1, -10.0, 10.0, 1e-4, 1e-4, 1, None, None)
    z = C.c_int64()
    with pytest.raises(N.FlowMCNativeError, match="n_dim"):
        N.call("fmc_flow_sizes", C.byref(f), C.byref(z), None, None, None, None, None)
    f = N.fmc_flow(4, 2, 8, 32, 32, 7, -10.0, 10.0, 1e-4, 1e-4, 1, None, None)
    with pytest.raises(N.FlowMCNativeError, match="precision"):
        N.call("fmc_flow_sizes", C.byref(f), C.byref(z), None, None, None, None, None)
    f = N.fmc_flow(4, 2, 8, 32, 32, 1, -10.0, 10.0, 1e-4, 1e-4, 1, None, None)
    with pytest.raises(N.FlowMCNativeError, match="packed buffers"):
        N.call("fmc_flow_pass", None, C.byref(f), 0, 0, 1, 4, 0, 0, 0, None, None, None, None, None)
    # the base log-prob of a reference-order inverse pass is refused (its output
    # does not have the forward map's density); checked before any launch
    f = N.fmc_flow(4, 2, 8, 32, 32, 1, -10.0, 10.0, 1e-4, 1e-4, 8, 16, 24)
    with pytest.raises(N.FlowMCNativeError, match="base log-prob"):
        N.call("fmc_flow_pass", None, C.byref(f), 1, 2 | 4, 0, 4, 0, 0, 0, None, None, None, None, None)
