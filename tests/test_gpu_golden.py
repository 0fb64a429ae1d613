"""Kernels against the committed golden vectors (tests/golden/*.npz).

Tolerances as in test_gpu_chains.py / test_gpu_flow.py (FP32 kernels vs
FP64 vectors): 2e-4 scaled error; MH decisions identical away from the
threshold (margin 2e-4 on the FP64 log-ratio)."""

import ctypes as C
import os

import numpy as np
import pytest
import torch

import flowmc_amd._native as N
from flowmc_amd import random as fr
from flowmc_amd import targets as TG
from flowmc_amd.resource.kernel.base import StepOutputs
from flowmc_amd.resource.kernel.HMC import HMC
from flowmc_amd.resource.kernel.MALA import MALA
from flowmc_amd.resource.logPDF import LogPDF
from flowmc_amd.resource.model.nf_model.rqSpline import MaskedCouplingRQSpline

pytestmark = pytest.mark.gpu
HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return {k: torch.from_numpy(v) for k, v in np.load(os.path.join(HERE, f"{name}.npz")).items()}


def close(a, b, tol=2e-4):
    a = a.detach().cpu().double()
    b = b.double()
    return bool(((a - b).abs() <= tol * (1 + b.abs())).all())


def test_golden_spline_through_a_one_layer_flow(gpu):
    gv = load("spline")
    K = int(gv["K"])
    P = 3 * K + 1
    for i in range(0, 64, 8):
        m = MaskedCouplingRQSpline(2, 1, [32, 32], K, fr.PRNGKey(0), precision="float32")
        with torch.no_grad():
            m.view(0, "W3").zero_()
            b3 = m.view(0, "b3")
            b3.zero_()
            b3[0:P] = gv["raw"][i].float().cuda()  # layer 0 transforms dim 0
            m.mark_updated()
        x = torch.tensor([[float(gv["x"][i]), 0.0]], device=gpu)
        y, ld = m.forward(x)
        assert close(y[:, 0], gv["y"][i:i + 1]) and close(ld, gv["ld"][i:i + 1])
        xi, ldi = m.inverse(x)
        assert close(xi[:, 0], gv["x_inv"][i:i + 1]) and close(ldi, gv["ld_inv"][i:i + 1])


def test_golden_flow(gpu):
    gv = load("flow")
    d, L, K, h1, h2 = [int(v) for v in gv["dims"]]
    m = MaskedCouplingRQSpline(d, L, [h1, h2], K, fr.PRNGKey(0), precision="float32", _theta=gv["theta"].float())
    x = gv["x"].float().cuda()
    assert close(m.log_prob(x), gv["log_prob"])
    y, ld = m.forward(x)
    assert close(y, gv["y"]) and close(ld, gv["ld"])
    xi, ldi = m.inverse(x)
    assert close(xi, gv["x_inv"]) and close(ldi, gv["ld_inv"])
    assert close(m.sample_from_noise(gv["z"].float().cuda()), gv["sample"])
    gt, _ = m.grad_log_prob_sum(x, torch.full((x.shape[0],), -1.0 / x.shape[0]).cuda())
    n = L * m._layer_size
    ref = gv["loss_grad"][:n]
    assert float((gt.cpu().double()[:n] - ref).abs().max()) <= 2e-3 * float(ref.abs().max())


def _window(kind, gv, prop, logpdf):
    x0 = gv["x0"].float().cuda()
    S, C_ = gv["log_u"].shape
    d = x0.shape[1]
    pos = torch.zeros(C_, S, d, device="cuda")
    lp = torch.zeros(C_, S, device="cuda")
    ac = torch.zeros(C_, S, device="cuda")
    prop.sample_steps(fr.PRNGKey(0), x0, logpdf, None, S, 1, StepOutputs(pos, lp, ac, 0),
                      noise=gv["noise"].float().cuda(), log_u=gv["log_u"].float().cuda())
    ok = ~((gv["ratios"] - gv["log_u"].T).abs() < 2e-4).any(1)
    assert int(ok.sum()) >= 0.85 * C_
    assert torch.equal(ac.cpu().double()[ok], gv["accepts"][ok])
    assert close(pos[ok.cuda()], gv["positions"][ok]) and close(lp[ok.cuda()], gv["log_probs"][ok])


def test_golden_mala_window(gpu):
    gv = load("mala")
    tgt = TG.GaussianDense(gv["mean"].float(), gv["precision"])
    _window("mala", gv, MALA(float(gv["step_size"]), fused_steps=32), LogPDF(tgt, n_dims=tgt.n_dim))


def test_golden_hmc_window(gpu):
    gv = load("hmc")
    d = gv["x0"].shape[1]
    _window("hmc", gv, HMC(gv["metric"].float(), float(gv["step_size"]), int(gv["n_leapfrog"]), fused_steps=7),
            LogPDF(TG.Funnel(d), n_dims=d))


def test_golden_nf_mh(gpu):
    gv = load("nf_mh")
    C_, S, d = gv["prop"].shape
    dev = [gv[k].float().contiguous().cuda() for k in ("x0", "lp0", "q0", "prop", "lp_prop", "q_prop", "log_u")]
    pos = torch.zeros(C_, S, d, device=gpu)
    lps = torch.zeros(C_, S, device=gpu)
    acs = torch.zeros(C_, S, device=gpu)
    idx = torch.empty(C_, S, dtype=torch.int32, device=gpu)
    N.call("fmc_nf_mh", N.stream_ptr(), *[t.data_ptr() for t in dev[:6]], C_, S, d, 0, 0, dev[6].data_ptr(),
           idx.data_ptr(), None, None, pos.data_ptr(), lps.data_ptr(), acs.data_ptr(), S, 0, 1)
    assert torch.equal(acs.cpu().bool(), gv["accepts"].bool())
    assert close(pos, gv["positions"], 1e-7) and close(lps, gv["log_probs"], 1e-7)


def test_golden_adamw(gpu):
    gv = load("adamw")
    n = gv["theta"].numel()
    th = gv["theta"].float().cuda()
    m = torch.zeros(n, device=gpu)
    v = torch.zeros(n, device=gpu)
    ws = torch.empty(1024, dtype=torch.float64, device=gpu)
    for t in range(3):
        g = gv["grads"][t].float().cuda()
        N.call("fmc_adamw_clip", N.stream_ptr(), th.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), n,
               float(gv["lr"]), 0.9, 0.999, 1e-8, 1e-4, 1.0, t + 1, ws.data_ptr(), 1024, None)
        assert float((th.cpu().double() - gv["thetas"][t]).abs().max()) <= 1e-5
