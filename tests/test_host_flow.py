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
    # compact conditioner: only conditioning columns of W1 and transformed rows of W3 are packed
    th, pk, _, _, _, _ = sizes(64, 2, 8, 256, 256)
    full = 2 * (2 * (256 * 64 + 256 * 256 + 64 * 25 * 256))  # forward + transposed images at full width
    assert pk < 0.75 * full


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
