"""TrainModel's training-data selection (src/flowMC/strategy/train_model.py:66-80)
against a direct restatement: boolean-mask the finite rows of the whole buffer,
reshape to (n_chains, -1, n_dims), slice [:, -history_window:], index with the
drawn rows. Bit-exact (the selection only moves rows)."""

import pytest
import torch

import flowmc_amd._native as N
from flowmc_amd import random as fr
from flowmc_amd.strategy.train_model import select_training_data

pytestmark = pytest.mark.gpu


def direct(data, history_window, n_max_examples, key):
    n_chains, n_dims = data.shape[0], data.shape[-1]
    rows = data[torch.isfinite(data).all(-1)]
    if rows.shape[0] % n_chains != 0:
        raise ValueError("different numbers of finite rows")
    rows = rows.reshape(n_chains, -1, n_dims)[:, -history_window:].reshape(-1, n_dims)
    if rows.shape[0] == 0:
        raise ValueError("no finite rows")
    idx = fr.choice(key, rows.shape[0], n_max_examples, device=rows.device)
    return rows[idx]


@pytest.mark.parametrize("d", [1, 3, 5, 64, 130])
def test_finite_rows_kernel(gpu, d):
    g = torch.Generator(device="cuda").manual_seed(d)
    x = torch.randn(3001, d + 1, device="cuda", generator=g)
    bad = torch.rand(3001, d + 1, device="cuda", generator=g)
    x[bad < 0.001] = float("inf")
    x[(bad > 0.5) & (bad < 0.5005)] = float("nan")
    x[bad > 0.9995] = -float("inf")
    for view in (x[:, :d].contiguous(), x[:, 1:].contiguous().view(-1)[1:1 + 3000 * d].view(3000, d)):
        out = torch.full((view.shape[0],), 7, dtype=torch.int32, device="cuda")
        N.call("fmc_finite_rows", N.stream_ptr(), view.data_ptr(), view.shape[0], d, out.data_ptr())
        torch.cuda.synchronize()
        assert torch.equal(out.bool(), torch.isfinite(view).all(-1)), d


@pytest.mark.parametrize("d", [5, 64])
def test_finite_rows_kernel_many_passes(gpu, d):
    """More rows than one launch covers in a pass (2048 workgroups x 256 / g
    rows, four passes per loop iteration), with a ragged last iteration."""
    n = 2048 * 64 * 4 * 3 + 1237 if d == 64 else 2048 * 32 * 4 + 999
    g = torch.Generator(device="cuda").manual_seed(100 + d)
    x = torch.randn(n, d, device="cuda", generator=g)
    bad_rows = torch.randint(0, n, (500,), device="cuda", generator=g)
    bad_cols = torch.randint(0, d, (500,), device="cuda", generator=g)
    x[bad_rows, bad_cols] = float("nan")
    x[::7, 0] = -float("inf")  # decided by a row's first four entries
    x[3::11] = -float("inf")  # whole rows, as in an unfilled sample buffer
    x[-1, d - 1] = float("inf")
    out = torch.full((n,), 7, dtype=torch.int32, device="cuda")
    N.call("fmc_finite_rows", N.stream_ptr(), x.data_ptr(), n, d, out.data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(out.bool(), torch.isfinite(x).all(-1))


def _buffer(C_, T, d, filled, seed=0):
    g = torch.Generator(device="cuda").manual_seed(seed)
    buf = torch.full((C_, T, d), -float("inf"), device="cuda")
    buf[:, :filled] = torch.randn(C_, filled, d, device="cuda", generator=g)
    return buf


@pytest.mark.parametrize("window", [7, 50, 200, 0, -3])
def test_selection_matches_direct(gpu, window):
    buf = _buffer(64, 120, 5, filled=37)
    key = fr.PRNGKey(11)
    a = select_training_data(buf, window, 1000, key)
    b = direct(buf, window, 1000, key)
    assert torch.equal(a, b)


def test_selection_keeps_the_reference_reshape_of_uneven_chains(gpu):
    # chain 0 has 4 finite rows and chain 1 has 2: 6 rows reshape to (2, 3, d)
    # across the chain boundary, as boolean indexing + reshape does
    buf = _buffer(2, 6, 3, filled=4, seed=1)
    buf[1, 2:4] = float("nan")
    key = fr.PRNGKey(2)
    for w in (2, 3, 5):
        assert torch.equal(select_training_data(buf, w, 64, key), direct(buf, w, 64, key))
    buf[1, 1] = float("inf")  # 5 finite rows over 2 chains
    with pytest.raises(ValueError):
        select_training_data(buf, 2, 8, key)
    with pytest.raises(ValueError):
        select_training_data(torch.full((3, 4, 2), -float("inf"), device="cuda"), 2, 8, key)


def test_selection_on_a_large_buffer(gpu):
    # the shape of the NF-MCMC training buffer (4096 chains x 4000 rows x 64) is
    # too large to filter directly in a test; 512 x 4000 x 64 keeps the layout
    buf = _buffer(512, 4000, 64, filled=1800, seed=3)
    key = fr.PRNGKey(5)
    assert torch.equal(select_training_data(buf, 200, 100000, key), direct(buf, 200, 100000, key))
