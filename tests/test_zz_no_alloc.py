"""fmc_nf_mh_noalloc (no device allocation: the proposal indices live in the
library's fixed per-device array) against fmc_nf_mh (caller-provided index
buffer) through the C ABI: bit-identical buffers, final states and
log-probabilities on two calls in a row, the second smaller than the first,
and the scan itself against the oracle's NF_proposal.py:91-126 restatement."""

import numpy as np
import pytest
import torch

import flowmc_amd._native as N
from oracle import chains as OC

pytestmark = pytest.mark.gpu


def _case(C_, S, d, seed, dev):
    g = torch.Generator().manual_seed(seed)
    x0 = torch.randn(C_, d, generator=g)
    prop = torch.randn(C_, S, d, generator=g)
    lp0, q0 = torch.randn(C_, generator=g), torch.randn(C_, generator=g)
    lpp, qp = torch.randn(C_, S, generator=g), torch.randn(C_, S, generator=g)
    lu = torch.log(torch.rand(C_, S, generator=g))
    return [t.to(dev).contiguous() for t in (x0, lp0, q0, prop, lpp, qp, lu)]


def _run(noalloc, case, thinning, buf_len, cursor, injected):
    x0, lp0, q0, prop, lpp, qp, lu = case
    C_, S, d = prop.shape
    dev = x0.device
    out_pos = torch.full((C_, buf_len, d), float("nan"), device=dev)
    out_lp = torch.full((C_, buf_len), float("nan"), device=dev)
    out_acc = torch.full((C_, buf_len), float("nan"), device=dev)
    x_last = torch.full_like(x0, float("nan"))
    lp_last = torch.full_like(lp0, float("nan"))
    common = (x0.data_ptr(), lp0.data_ptr(), q0.data_ptr(), prop.data_ptr(), lpp.data_ptr(), qp.data_ptr(), C_, S, d,
              123, 456, lu.data_ptr() if injected else None)
    tail = (x_last.data_ptr(), lp_last.data_ptr(), out_pos.data_ptr(), out_lp.data_ptr(), out_acc.data_ptr(), buf_len,
            cursor, thinning)
    if noalloc:
        N.call("fmc_nf_mh_noalloc", N.stream_ptr(), *common, *tail)
    else:
        idx = torch.empty(C_, S, dtype=torch.int32, device=dev)
        N.call("fmc_nf_mh", N.stream_ptr(), *common, idx.data_ptr(), *tail)
    torch.cuda.synchronize()
    return [t.cpu() for t in (out_pos, out_lp, out_acc, x_last, lp_last)]


@pytest.mark.parametrize("injected", [True, False])
def test_noalloc_matches_the_allocating_entry_point(gpu, injected):
    # two calls in a row on the shared array, the second smaller
    for (C_, S, d, thinning, seed) in [(4096, 100, 64, 1, 0), (300, 37, 5, 3, 1)]:
        case = _case(C_, S, d, seed, gpu)
        n_rows = (S + thinning - 1) // thinning
        want = _run(False, case, thinning, n_rows + 4, 2, injected)
        got = _run(True, case, thinning, n_rows + 4, 2, injected)
        for w, g_ in zip(want, got):
            assert torch.equal(torch.nan_to_num(w, nan=7.5), torch.nan_to_num(g_, nan=7.5))
        if injected:
            # the scan against the oracle on the same log-uniforms, run in FP32:
            # the ratio (lp' - lp) - (q' - q) is two IEEE subtractions either
            # way, so decisions and rows agree bit for bit
            x0, lp0, q0, prop, lpp, qp, lu = [t.cpu() for t in case]
            pos, lps, accs = OC.nf_mh_scan(x0, lp0, q0, prop, lpp, qp, lu)
            rows = slice(2, 2 + n_rows)
            assert torch.equal(got[0][:, rows], pos[:, ::thinning])
            assert torch.equal(got[1][:, rows], lps[:, ::thinning])
            assert torch.equal(got[2][:, rows], accs[:, ::thinning].float())
            assert torch.equal(got[3], pos[:, ((S - 1) // thinning) * thinning])


def test_noalloc_refuses_more_than_its_array_holds(gpu):
    lib = N.load()
    S = 100
    C_ = N.NF_MH_NOALLOC_MAX // S + 1
    rc = lib.fmc_nf_mh_noalloc(None, 1, 1, 1, 1, 1, 1, C_, S, 4, 0, 0, None, None, None, None, None, None, 1, 0, 1)
    assert rc == 1 and "exceeds" in N.last_error()
