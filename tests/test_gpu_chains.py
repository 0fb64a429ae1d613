"""GPU parity of the chain kernels against the FP64 oracle.

Tolerances (FP32 kernels vs FP64 oracle on identical injected noise):
  * target values / gradients: |err| <= 2e-5 * (1 + |ref|) (relative to scale)
  * MALA / HMC trajectories over the 100-step window: positions within
    atol 2e-4 * (1 + |x|); log-probs within 2e-4 * (1 + |lp|)
  * accept/reject decisions identical, except on chains where some step's
    oracle log-ratio lies within MARGIN of the log-uniform (counted; such a
    chain is excluded from the position comparison and must be rare).
"""

import ctypes as C

import numpy as np
import pytest
import torch

import flowmc_amd._native as N
from flowmc_amd import random as fr
from flowmc_amd import targets as TG
from flowmc_amd.resource.buffers import Buffer
from flowmc_amd.resource.kernel.base import StepOutputs
from flowmc_amd.resource.kernel.Gaussian_random_walk import GaussianRandomWalk
from flowmc_amd.resource.kernel.HMC import HMC
from flowmc_amd.resource.kernel.MALA import MALA
from flowmc_amd.resource.logPDF import LogPDF
from flowmc_amd.resource.states import State
from flowmc_amd.strategy.take_steps import TakeSerialSteps
from oracle import chains as OC
from oracle import philox as OP
from oracle import targets as OT

pytestmark = pytest.mark.gpu

MARGIN = 2e-4


def make_target(name, d, device="cuda", seed=0):
    g = torch.Generator().manual_seed(seed)
    if name == "gauss_diag":
        mean = torch.randn(d, generator=g, dtype=torch.float64)
        prec = 0.5 + torch.rand(d, generator=g, dtype=torch.float64)
        return TG.GaussianDiag(mean, prec, device), lambda z: OT.gauss_diag(z, mean, prec)
    if name == "gauss_dense":
        t = TG.correlated_gaussian_5d(seed, n_dim=d, device=device)
        P = t.precision.to(torch.float64)
        return t, lambda z: OT.gauss_dense(z, t.mean.to(torch.float64), P)
    if name == "rosenbrock":
        return TG.Rosenbrock(d, device=device), lambda z: OT.rosenbrock(z)
    if name == "funnel":
        return TG.Funnel(d, device=device), OT.funnel
    if name == "gmm":
        t = TG.gaussian_mixture_64d(seed, n_dim=d, n_modes=4, device=device)
        return t, lambda z: OT.gmm_diag(z, t.means, t.sigmas)
    if name == "gmm_w":  # unequal weights and per-mode / per-dim scales
        means = 2.5 * torch.randn(5, d, generator=g, dtype=torch.float64)
        sig = 0.4 + 1.2 * torch.rand(5, d, generator=g, dtype=torch.float64)
        w = torch.tensor([0.05, 0.4, 0.1, 0.3, 0.15], dtype=torch.float64)
        t = TG.GaussianMixtureDiag(means, sig, w, device=device)
        return t, lambda z: OT.gmm_diag(z, means, sig, torch.log(w))
    if name == "dual_moon":
        c = torch.arange(d, dtype=torch.float64) * 0.1
        return TG.DualMoon(d, c, device), lambda z: OT.dual_moon(z, c)
    raise ValueError(name)


def test_philox_blocks_bit_exact(gpu):
    key = (0xDEADBEEF, 0x12345678)
    n = 4096
    out = torch.empty(4 * n, dtype=torch.int32, device=gpu)
    N.call("fmc_philox_blocks", N.stream_ptr(), key[0], key[1], 7, 3, 11, 2, n, out.data_ptr())
    got = out.cpu().numpy().view(np.uint32).reshape(n, 4)
    ctr = np.stack([7 + np.arange(n), np.full(n, 3), np.full(n, 11), np.full(n, 2)], -1).astype(np.uint32)
    want = OP.philox4x32_10(ctr, key)
    assert np.array_equal(got, want)


def test_rng_streams_match_oracle(gpu):
    key = (5, 77)
    C_, d, S = 300, 13, 3
    z = torch.empty(S, C_, d, device=gpu, dtype=torch.float32)
    N.call("fmc_rng_normal", N.stream_ptr(), key[0], key[1], C_, d, S, 10, OP.RNG_MALA_NOISE, z.data_ptr())
    want = OP.normals(key, C_, d, S, 10, OP.RNG_MALA_NOISE)
    assert np.allclose(z.cpu().numpy(), want, atol=2e-6 * (1 + np.abs(want)).max(), rtol=2e-6)
    lu = torch.empty(S, C_, device=gpu, dtype=torch.float32)
    N.call("fmc_rng_log_uniform", N.stream_ptr(), key[0], key[1], C_, S, 10, OP.RNG_MALA_UNIFORM, lu.data_ptr())
    want = OP.log_uniforms(key, C_, S, 10, OP.RNG_MALA_UNIFORM)
    assert np.allclose(lu.cpu().numpy(), want, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("name,d", [("gauss_diag", 5), ("gauss_diag", 130), ("gauss_dense", 5), ("gauss_dense", 12),
                                    ("rosenbrock", 128), ("rosenbrock", 7), ("funnel", 16), ("funnel", 201),
                                    ("gmm", 64), ("gmm", 9), ("gmm_w", 24), ("dual_moon", 5)])
@pytest.mark.parametrize("layout", [None, (1, 8), (4, 2), (64, 1)])
def test_target_value_and_grad(gpu, name, d, layout):
    tgt, fn = make_target(name, d)
    cfg = None
    if layout is not None:
        lpc, nv = layout
        if 4 * lpc * nv < d:
            pytest.skip("layout too small for d")
        cfg = N.fmc_launch_cfg(lpc, nv, 256, 0)
    g = torch.Generator().manual_seed(1)
    x64 = torch.randn(257, d, generator=g, dtype=torch.float64) * 0.8
    if name == "funnel":
        x64[:, 0] = x64[:, 0].clamp(-2, 2)
    x = x64.to(torch.float32).to(gpu)
    lp = torch.empty(257, device=gpu, dtype=torch.float32)
    gr = torch.empty_like(x)
    desc = tgt.descriptor()
    N.call("fmc_target_eval", N.stream_ptr(), C.byref(desc), x.data_ptr(), 257, lp.data_ptr(), gr.data_ptr(),
           C.byref(cfg) if cfg is not None else None)
    v_ref, g_ref = OT.value_and_grad(fn, x.cpu().double())
    scale_v = 1 + v_ref.abs()
    assert ((lp.cpu().double() - v_ref).abs() <= 2e-5 * scale_v).all()
    scale_g = 1 + g_ref.abs().amax(-1, keepdim=True)
    # The mixture gradient sum_k w_k g_k weighs per-mode gradients by softmax(q): an FP32
    # error u*|q_k| in the mode log-densities (|q_k| ~ d/2 + |c_k| ~ 40 at d = 64) moves the
    # weights by that much relative, and the gradient by that times the spread of the g_k
    # (~10 here): ~1e-4 relative to max|g| where two modes compete. Other targets: 2e-5.
    tol_g = 1e-4 if name in ("gmm", "gmm_w") else 2e-5
    assert ((gr.cpu().double() - g_ref).abs() <= tol_g * scale_g).all()


def dense_condition_matrix(d, g):
    """An SPD condition matrix with off-diagonal couplings and positive
    column sums (the kinetic weights of the reference's broadcast kinetic)."""
    A = torch.randn(d, d, generator=g, dtype=torch.float64)
    S = (A @ A.T) / d
    M = torch.diag(0.5 + torch.rand(d, generator=g, dtype=torch.float64)) + 0.2 * S + 0.05
    return M


def _compare_traj(kind, name, d, C_, S, dt, thinning=1, n_leapfrog=5, seed=0, dense=False, target=None):
    tgt, fn = make_target(name, d, seed=seed) if target is None else target
    g = torch.Generator().manual_seed(seed + 11)
    x0 = torch.randn(C_, d, generator=g, dtype=torch.float64) * 0.5
    noise = torch.randn(S, C_, d, generator=g, dtype=torch.float64)
    log_u = torch.log(torch.rand(S, C_, generator=g, dtype=torch.float64))
    metric = 0.5 + torch.rand(d, generator=g, dtype=torch.float64) if kind == "hmc" else None
    if dense:
        metric = dense_condition_matrix(d, g)
        assert bool((metric.sum(0) > 0).all())
    n_rows = (S + thinning - 1) // thinning
    pos = torch.full((C_, n_rows + 3, d), float("nan"), device="cuda", dtype=torch.float32)
    lps = torch.full((C_, n_rows + 3), float("nan"), device="cuda", dtype=torch.float32)
    accs = torch.full((C_, n_rows + 3), float("nan"), device="cuda", dtype=torch.float32)
    out = StepOutputs(pos, lps, accs, 2)
    if kind == "mala":
        prop = MALA(dt, fused_steps=32)
    elif kind == "grw":
        prop = GaussianRandomWalk(dt, fused_steps=32)
    else:
        prop = HMC(metric.to(torch.float32), dt, n_leapfrog, fused_steps=8)
    logpdf = LogPDF(tgt, n_dims=d)
    last = prop.sample_steps(fr.PRNGKey(0), x0.float().cuda(), logpdf, None, S, thinning, out,
                             noise=noise.float().cuda(), log_u=log_u.float().cuda())
    torch.cuda.synchronize()
    # oracle on the FP32-rounded inputs
    kw = dict(step_size=dt)
    if kind == "hmc":
        kw.update(n_leapfrog=n_leapfrog, metric=metric.float().double())
    r_pos, r_lp, r_acc, r_last, ratios = OC.take_serial_steps(kind, x0.float().double(), fn, noise.float().double(),
                                                              log_u.float().double(), thinning, **kw)
    close = ((ratios - log_u.float().double().T).abs() < MARGIN).any(1)  # [C]
    n_close = int(close.sum())
    assert n_close <= max(2, (15 * C_) // 100), f"{n_close} chains near the MH threshold"
    ok = ~close
    gp = pos[:, 2:2 + n_rows].cpu().double()
    gl = lps[:, 2:2 + n_rows].cpu().double()
    ga = accs[:, 2:2 + n_rows].cpu().double()
    assert torch.equal(ga[ok], r_acc[ok]), "accept/reject decisions differ"
    assert ((gp[ok] - r_pos[ok]).abs() <= 2e-4 * (1 + r_pos[ok].abs())).all()
    assert ((gl[ok] - r_lp[ok]).abs() <= 2e-4 * (1 + r_lp[ok].abs())).all()
    assert ((last.cpu().double()[ok] - r_last[ok]).abs() <= 2e-4 * (1 + r_last[ok].abs())).all()
    # rows outside the written window are untouched
    assert torch.isnan(pos[:, :2]).all() and torch.isnan(pos[:, 2 + n_rows:]).all()
    return ga


@pytest.mark.parametrize("name,d", [("gauss_dense", 5), ("rosenbrock", 128), ("gmm", 64), ("funnel", 20),
                                    ("dual_moon", 5), ("gauss_diag", 33)])
def test_mala_fused_matches_oracle(gpu, name, d):
    dt = {"gauss_dense": 0.3, "rosenbrock": 0.01, "gmm": 0.2, "funnel": 0.2, "dual_moon": 0.05,
          "gauss_diag": 0.4}[name]
    acc = _compare_traj("mala", name, d, 192, 100, dt)
    assert 0.05 < float(acc.mean()) < 1.0


def test_mala_thinning_and_cursor(gpu):
    _compare_traj("mala", "gauss_diag", 6, 64, 23, 0.5, thinning=4)


@pytest.mark.parametrize("name,d", [("gauss_diag", 8), ("funnel", 16), ("rosenbrock", 12), ("gmm", 10)])
def test_hmc_fused_matches_oracle(gpu, name, d):
    dt = {"gauss_diag": 0.2, "funnel": 0.05, "rosenbrock": 0.02, "gmm": 0.1}[name]
    acc = _compare_traj("hmc", name, d, 128, 40, dt, n_leapfrog=6)
    assert float(acc.mean()) > 0.05


@pytest.mark.parametrize("name,d", [("gauss_dense", 5), ("rosenbrock", 128), ("gmm", 64), ("funnel", 20),
                                    ("gauss_diag", 33)])
def test_grw_fused_matches_oracle(gpu, name, d):
    dt = {"gauss_dense": 0.5, "rosenbrock": 0.02, "gmm": 0.1, "funnel": 0.2, "gauss_diag": 0.3}[name]
    acc = _compare_traj("grw", name, d, 192, 100, dt)
    assert 0.05 < float(acc.mean()) < 1.0


def test_grw_thinning_and_cursor(gpu):
    _compare_traj("grw", "gauss_diag", 6, 64, 23, 0.5, thinning=4)


def test_grw_generic_target_matches_oracle(gpu):
    # user-supplied per-point torch callable: propose / accept kernels around a batched evaluation
    d, C_, S, dt = 4, 96, 40, 0.6
    g = torch.Generator().manual_seed(21)
    x0 = torch.randn(C_, d, generator=g)
    noise = torch.randn(S, C_, d, generator=g)
    log_u = torch.log(torch.rand(S, C_, generator=g))
    mean = torch.tensor([0.5, -1.0, 0.0, 2.0])

    def user_logpdf(x, data):
        return -0.5 * torch.sum((x - data["mu"]) ** 2 / torch.arange(1, 5, device=x.device, dtype=x.dtype))

    pos = torch.zeros(C_, S, d, device=gpu)
    lps = torch.zeros(C_, S, device=gpu)
    accs = torch.zeros(C_, S, device=gpu)
    GaussianRandomWalk(dt).sample_steps(fr.PRNGKey(0), x0.cuda(), LogPDF(user_logpdf, n_dims=d),
                                        {"mu": mean.cuda()}, S, 1, StepOutputs(pos, lps, accs, 0),
                                        noise=noise.cuda(), log_u=log_u.cuda())
    fn = lambda z: -0.5 * (((z - mean.double()) ** 2) / torch.arange(1, 5, dtype=torch.float64)).sum(-1)
    r_pos, r_lp, r_acc, _, ratios = OC.take_serial_steps("grw", x0.double(), fn, noise.double(), log_u.double(),
                                                         step_size=dt)
    ok = ~((ratios - log_u.double().T).abs() < MARGIN).any(1)
    assert int((~ok).sum()) <= 3
    assert torch.equal(accs.cpu().double()[ok], r_acc[ok])
    assert ((pos.cpu().double()[ok] - r_pos[ok]).abs() <= 2e-4 * (1 + r_pos[ok].abs())).all()
    assert ((lps.cpu().double()[ok] - r_lp[ok]).abs() <= 2e-4 * (1 + r_lp[ok].abs())).all()


def test_grw_philox_mode_equals_injected_mode(gpu):
    tgt, _ = make_target("gmm", 16)
    C_, S, d = 300, 24, 16
    key = fr.PRNGKey(98)
    k0, k1 = fr.key_words(key)
    x0 = torch.randn(C_, d, device=gpu)
    noise = torch.empty(S, C_, d, device=gpu, dtype=torch.float32)
    lu = torch.empty(S, C_, device=gpu, dtype=torch.float32)
    N.call("fmc_rng_normal", N.stream_ptr(), k0, k1, C_, d, S, 0, OP.RNG_GRW_NOISE, noise.data_ptr())
    N.call("fmc_rng_log_uniform", N.stream_ptr(), k0, k1, C_, S, 0, OP.RNG_GRW_UNIFORM, lu.data_ptr())
    res = []
    for nz, l in ((None, None), (noise, lu)):
        pos = torch.zeros(C_, S, d, device=gpu, dtype=torch.float32)
        lp = torch.zeros(C_, S, device=gpu, dtype=torch.float32)
        ac = torch.zeros(C_, S, device=gpu, dtype=torch.float32)
        GaussianRandomWalk(0.2, fused_steps=8).sample_steps(key, x0, LogPDF(tgt, n_dims=d), None, S, 1,
                                                            StepOutputs(pos, lp, ac, 0), noise=nz, log_u=l)
        res.append((pos, lp, ac))
    for a, b in zip(*res):
        assert torch.equal(a, b)


def test_philox_mode_equals_injected_mode(gpu):
    tgt, _ = make_target("rosenbrock", 64)
    C_, S, d = 500, 40, 64
    key = fr.PRNGKey(99)
    k0, k1 = fr.key_words(key)
    x0 = torch.randn(C_, d, device=gpu) * 0.3
    noise = torch.empty(S, C_, d, device=gpu, dtype=torch.float32)
    lu = torch.empty(S, C_, device=gpu, dtype=torch.float32)
    N.call("fmc_rng_normal", N.stream_ptr(), k0, k1, C_, d, S, 0, OP.RNG_MALA_NOISE, noise.data_ptr())
    N.call("fmc_rng_log_uniform", N.stream_ptr(), k0, k1, C_, S, 0, OP.RNG_MALA_UNIFORM, lu.data_ptr())
    res = []
    for nz, l in ((None, None), (noise, lu)):
        pos = torch.zeros(C_, S, d, device=gpu, dtype=torch.float32)
        lp = torch.zeros(C_, S, device=gpu, dtype=torch.float32)
        ac = torch.zeros(C_, S, device=gpu, dtype=torch.float32)
        MALA(0.01, fused_steps=16).sample_steps(key, x0, LogPDF(tgt, n_dims=d), None, S, 1,
                                               StepOutputs(pos, lp, ac, 0), noise=nz, log_u=l)
        res.append((pos, lp, ac))
    for a, b in zip(*res):
        assert torch.equal(a, b)


def test_generic_target_path_matches_fused(gpu):
    # user-supplied torch callable (torch.func gradient + propose/accept kernels)
    d, C_, S = 5, 64, 30
    mean = torch.arange(d, dtype=torch.float32, device=gpu)

    def user_logpdf(x, data):
        return -0.5 * torch.sum((x - data["data"]) ** 2)

    native = LogPDF(TG.GaussianDiag(mean.cpu(), None), n_dims=d)
    generic = LogPDF(user_logpdf, n_dims=d)
    key = fr.PRNGKey(3)
    x0 = torch.randn(C_, d, device=gpu)
    outs = []
    for lpdf in (native, generic):
        pos = torch.zeros(C_, S, d, device=gpu, dtype=torch.float32)
        lp = torch.zeros(C_, S, device=gpu, dtype=torch.float32)
        ac = torch.zeros(C_, S, device=gpu, dtype=torch.float32)
        MALA(0.4).sample_steps(key, x0, lpdf, {"data": mean}, S, 1, StepOutputs(pos, lp, ac, 0))
        outs.append((pos, lp, ac))
    (p1, l1, a1), (p2, l2, a2) = outs
    agree = (a1 == a2).all(1)
    assert float(agree.float().mean()) > 0.95
    assert torch.allclose(p1[agree], p2[agree], atol=1e-4, rtol=1e-4)


def test_nf_mh_matches_oracle(gpu):
    C_, S, d = 300, 50, 7
    g = torch.Generator().manual_seed(5)
    x0 = torch.randn(C_, d, generator=g)
    lp0, q0 = torch.randn(C_, generator=g), torch.randn(C_, generator=g)
    prop = torch.randn(C_, S, d, generator=g)
    lpp, qp = torch.randn(C_, S, generator=g), torch.randn(C_, S, generator=g)
    lu = torch.log(torch.rand(C_, S, generator=g))
    r_pos, r_lp, r_acc = OC.nf_mh_scan(x0.double(), lp0.double(), q0.double(), prop.double(), lpp.double(),
                                       qp.double(), lu.double())
    for thin in (1, 3):
        n_rows = (S + thin - 1) // thin
        pos = torch.zeros(C_, n_rows + 1, d, device=gpu, dtype=torch.float32)
        lps = torch.zeros(C_, n_rows + 1, device=gpu, dtype=torch.float32)
        acs = torch.zeros(C_, n_rows + 1, device=gpu, dtype=torch.float32)
        idx = torch.empty(C_, S, dtype=torch.int32, device=gpu)
        xl = torch.empty(C_, d, device=gpu, dtype=torch.float32)
        ll = torch.empty(C_, device=gpu, dtype=torch.float32)
        dev = [t.contiguous().cuda() for t in (x0, lp0, q0, prop, lpp, qp, lu)]  # keep alive across the call
        N.call("fmc_nf_mh", N.stream_ptr(), *[t.data_ptr() for t in dev[:6]], C_, S, d, 0, 0, dev[6].data_ptr(),
               idx.data_ptr(), xl.data_ptr(), ll.data_ptr(), pos.data_ptr(), lps.data_ptr(), acs.data_ptr(),
               n_rows + 1, 1, thin)
        torch.cuda.synchronize()
        sel = slice(None, None, thin)
        assert torch.equal(acs[:, 1:].cpu().bool(), r_acc[:, sel])
        assert torch.equal(pos[:, 1:].cpu(), r_pos[:, sel].float())
        assert torch.equal(lps[:, 1:].cpu(), r_lp[:, sel].float())
        assert torch.equal(xl.cpu(), r_pos[:, sel][:, -1].float())


# ---------------------------------------- reference-style statistical tests
def _strategy_run(kernel, logpdf, n_chains, n_steps, d, key=42, x0=None):
    positions = Buffer("positions", (n_chains, n_steps, d), 1)
    log_prob = Buffer("log_prob", (n_chains, n_steps), 1)
    acceptance = Buffer("acceptance", (n_chains, n_steps), 1)
    state = State({"positions": "positions", "log_prob": "log_prob", "acceptance": "acceptance"},
                  name="sampler_state")
    resources = {"positions": positions, "log_prob": log_prob, "acceptance": acceptance, "K": kernel,
                 "logpdf": logpdf, "sampler_state": state}
    strat = TakeSerialSteps("logpdf", kernel_name="K", state_name="sampler_state",
                            buffer_names=["positions", "log_prob", "acceptance"], n_steps=n_steps)
    k, sub = fr.split(fr.PRNGKey(key))
    x0 = fr.normal(sub, (n_chains, d)) if x0 is None else x0
    _, res, last = strat(k, resources, x0, {})
    return res, last, strat


def test_mala_close_gaussian(gpu):
    # test/unit/test_kernels.py:241-288: N(0, I_2), step 1, 1 chain x 50k steps
    logpdf = LogPDF(TG.GaussianDiag(torch.zeros(2, dtype=torch.float32)), n_dims=2)
    res, _, strat = _strategy_run(MALA(1.0), logpdf, 1, 50000, 2)
    data = res["positions"].data
    assert abs(float(data.mean())) < 3e-2 and abs(float(data.var()) - 1) < 3e-2
    assert strat.current_position == 50000


def test_hmc_close_gaussian(gpu):
    # test/unit/test_kernels.py:135-184: N(0, I_2), step 0.1, 5 leapfrog, 30k steps
    logpdf = LogPDF(TG.GaussianDiag(torch.zeros(2, dtype=torch.float32)), n_dims=2)
    res, _, _ = _strategy_run(HMC(torch.eye(2), 0.1, 5), logpdf, 1, 30000, 2)
    data = res["positions"].data
    assert abs(float(data.mean())) < 3e-2 and abs(float(data.var()) - 1) < 3e-2


def test_grw_close_gaussian(gpu):
    # test/unit/test_kernels.py:352-392: N(0, I), step 1; 64 chains x 20k steps
    logpdf = LogPDF(TG.GaussianDiag(torch.zeros(2, dtype=torch.float32)), n_dims=2)
    res, _, strat = _strategy_run(GaussianRandomWalk(1.0), logpdf, 64, 20000, 2)
    data = res["positions"].data
    assert abs(float(data.mean())) < 3e-2 and abs(float(data.var()) - 1) < 3e-2
    assert strat.current_position == 20000


def test_grw_kernel_api(gpu):
    # test_kernels.py:293-350: repr, deterministic kernel, acceptance -> 1 as the step -> 0
    logpdf = LogPDF(TG.GaussianDiag(torch.zeros(2, dtype=torch.float32)), n_dims=2)
    x = fr.normal(fr.PRNGKey(4), (1, 2))[0]
    k = GaussianRandomWalk(step_size=1)
    r1 = k.kernel(fr.PRNGKey(7), x, logpdf(x), logpdf, None)
    r2 = k.kernel(fr.PRNGKey(7), x, logpdf(x), logpdf, None)
    assert torch.equal(r1[0], r2[0]) and bool(r1[1] == r2[1]) and bool(r1[2] == r2[2])
    xs = fr.normal(fr.PRNGKey(1), (100, 2))
    _, _, acc = GaussianRandomWalk(1e-5).kernel(fr.PRNGKey(2), xs, logpdf(xs), logpdf, None)
    assert bool(acc.all())


def test_acceptance_goes_to_one(gpu):
    logpdf = LogPDF(TG.GaussianDiag(torch.zeros(2, dtype=torch.float32)), n_dims=2)
    x = fr.normal(fr.PRNGKey(1), (100, 2))
    _, _, acc = MALA(1e-5).kernel(fr.PRNGKey(2), x, None, logpdf, None)
    assert bool(acc.all())
    _, _, acc = HMC(torch.eye(2), 1e-7, 5).kernel(fr.PRNGKey(2), x, logpdf(x), logpdf, None)
    assert bool(acc.all())


def test_kernels_deterministic(gpu):
    logpdf = LogPDF(TG.GaussianDiag(torch.zeros(2, dtype=torch.float32)), n_dims=2)
    x = fr.normal(fr.PRNGKey(4), (1, 2))[0]
    for k in (MALA(1.0), HMC(torch.eye(2), 1.0, 5)):
        r1 = k.kernel(fr.PRNGKey(7), x, logpdf(x), logpdf, None)
        r2 = k.kernel(fr.PRNGKey(7), x, logpdf(x), logpdf, None)
        assert torch.equal(r1[0], r2[0]) and torch.equal(r1[1], r2[1]) and bool(r1[2] == r2[2])


def test_nonfinite_target_rejects(gpu):
    # NaN / -inf log-densities never accept (MALA.py:84 compares log u < NaN -> False)
    d = 3

    def bad(x, data):
        return torch.where(x[0] > 0.0, torch.tensor(float("nan"), device=x.device), -0.5 * torch.sum(x * x))

    x0 = -torch.ones(64, d, device=gpu, dtype=torch.float32) * 0.1
    pos = torch.zeros(64, 20, d, device=gpu, dtype=torch.float32)
    lp = torch.zeros(64, 20, device=gpu, dtype=torch.float32)
    ac = torch.zeros(64, 20, device=gpu, dtype=torch.float32)
    MALA(0.5).sample_steps(fr.PRNGKey(0), x0, LogPDF(bad, n_dims=d), None, 20, 1, StepOutputs(pos, lp, ac, 0))
    assert bool(torch.isfinite(pos).all()) and bool((pos[..., 0] <= 0).all())


def test_batched_hmc_matches_single_target_runs(gpu):
    from flowmc_amd.resource.kernel.batched_HMC import BatchedHMC

    dims = TG.funnel_dims(3, n_targets=6, lo=8, hi=256)
    targets = [TG.Funnel(d) for d in dims]
    C_, S = 64, 12
    g = torch.Generator().manual_seed(9)
    noise = [torch.randn(S, C_, d, generator=g).cuda() for d in dims]
    log_u = [torch.log(torch.rand(S, C_, generator=g)).cuda() for _ in dims]
    x0 = [torch.randn(C_, d, generator=g).cuda() * 0.5 for d in dims]
    bh = BatchedHMC(targets, C_, step_size=0.05, n_leapfrog=10)
    for x, x_init in zip(bh.x, x0):
        x.copy_(x_init)
    outs = [(torch.zeros(C_, S, d, device=gpu), torch.zeros(C_, S, device=gpu), torch.zeros(C_, S, device=gpu), 0)
            for d in dims]
    bh.run(fr.PRNGKey(0), S, outputs=outs, noise=noise, log_u=log_u, fused_steps=5)
    torch.cuda.synchronize()
    for i, (t, d) in enumerate(zip(targets, dims)):
        pos = torch.zeros(C_, S, d, device=gpu)
        lp = torch.zeros(C_, S, device=gpu)
        ac = torch.zeros(C_, S, device=gpu)
        last = HMC(1, 0.05, 10, fused_steps=5).sample_steps(fr.PRNGKey(0), x0[i], LogPDF(t, n_dims=d), None, S, 1,
                                                           StepOutputs(pos, lp, ac, 0), noise=noise[i],
                                                           log_u=log_u[i])
        assert torch.equal(outs[i][2], ac), f"slot {i} (d={d}) acceptance differs"
        assert torch.allclose(outs[i][0], pos, atol=1e-6, rtol=1e-6)
        assert torch.allclose(bh.x[i], last, atol=1e-6, rtol=1e-6)
    # the autotuner returns a valid layout and the tuned launch reproduces the default-layout run
    layout = bh.tune(fr.PRNGKey(1))
    assert layout[0] in (1, 2) and all(4 * layout[0] * l >= d for l, d in zip(layout[2], dims))


@pytest.mark.parametrize("entry", ["fmc_mala_run", "fmc_grw_run"])
def test_fused_launches_replay_in_a_hip_graph(gpu, entry):
    # A sequence of fused launches captured as one HIP graph replays to the eager
    # result bit for bit, replay after replay, with eager work and host
    # synchronisation between the replays (the launches record no memsets).
    C_, d, K, S = 1024, 128, 16, 64
    tgt = TG.Rosenbrock(d, device="cuda")
    desc = tgt.descriptor()
    x0 = fr.normal(fr.PRNGKey(11), (C_, d), device="cuda")
    x = x0.clone()
    lp = torch.empty(C_, device="cuda")
    pos = torch.empty((C_, S, d), device="cuda")
    lps = torch.empty((C_, S), device="cuda")
    acc = torch.empty((C_, S), device="cuda")
    k0, k1 = fr.key_words(fr.PRNGKey(12))

    def run():
        st = torch.cuda.current_stream().cuda_stream
        for s0 in range(0, S, K):
            N.call(entry, st, C.byref(desc), x.data_ptr(), lp.data_ptr(), C_, K, s0, 0.01, k0, k1, None, None,
                   pos.data_ptr(), lps.data_ptr(), acc.data_ptr(), S, 0, 1, None)

    run()
    torch.cuda.synchronize()
    want = [t.clone() for t in (x, lp, pos, lps, acc)]
    x.copy_(x0)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        run()
    torch.cuda.synchronize()
    for rep in range(3):
        x.copy_(x0)
        for t in (lp, pos, lps, acc):
            t.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = (x, lp, pos, lps, acc)
        assert all(torch.equal(a, b) for a, b in zip(got, want)), f"replay {rep} differs from the eager run"
        _ = (x0 * 2).sum().item()  # eager work between replays


def test_batched_hmc_replays_in_a_hip_graph(gpu):
    # BatchedHMC.prepare uploads the tables once; its launches recorded as one HIP
    # graph replay to the eager result bit for bit, replay after replay.
    from flowmc_amd.resource.kernel.batched_HMC import BatchedHMC

    dims = TG.funnel_dims(5, n_targets=5, lo=8, hi=200)
    C_, S = 96, 32
    bh = BatchedHMC([TG.Funnel(d) for d in dims], C_, step_size=0.05, n_leapfrog=10)
    bh.init_positions(fr.PRNGKey(3))
    x0 = [x.clone() for x in bh.x]
    outs = [(torch.empty(C_, S, d, device=gpu), torch.empty(C_, S, device=gpu), torch.empty(C_, S, device=gpu), 0)
            for d in dims]
    launch = bh.prepare(fr.PRNGKey(4), S, outputs=outs, fused_steps=8)
    launch()
    torch.cuda.synchronize()
    want = [t.clone() for o in outs for t in o[:3]] + [x.clone() for x in bh.x] + [lp.clone() for lp in bh.logp]
    for x, x_init in zip(bh.x, x0):
        x.copy_(x_init)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        launch()
    torch.cuda.synchronize()
    for rep in range(3):
        for x, x_init in zip(bh.x, x0):
            x.copy_(x_init)
        for o in outs:
            for t in o[:3]:
                t.fill_(float("nan"))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        got = [t for o in outs for t in o[:3]] + list(bh.x) + list(bh.logp)
        assert all(torch.equal(a, b) for a, b in zip(got, want)), f"replay {rep} differs from the eager run"
        _ = sum(float(x.sum()) for x in x0)  # eager work between replays


@pytest.mark.parametrize("name,d,L", [("funnel", 256, 10), ("gmm", 64, 10), ("gmm_w", 24, 10)])
def test_hmc_fused_matches_oracle_long_trajectories(gpu, name, d, L):
    # the funnel workload's trajectory (L = 10, eps = 0.05) at its largest dimension
    dt = {"funnel": 0.05, "gmm": 0.05, "gmm_w": 0.1}[name]
    acc = _compare_traj("hmc", name, d, 96, 12, dt, n_leapfrog=L)
    assert float(acc.mean()) > 0.05


@pytest.mark.parametrize("metric", ["mixed", "unit"])
def test_batched_hmc_matches_oracle(gpu, metric):
    # the uneven-dimension batch against the FP64 oracle, target by target:
    # 9 funnels with d from 8 to 256, L = 10, eps = 0.05, injected noise;
    # "mixed": unit metric on some slots and a diagonal metric on others (the
    # run-time dispatch kernel), "unit": every slot on the unit metric (the
    # kernel specialised on the funnel and on each slot's lanes per chain)
    from flowmc_amd.resource.kernel.batched_HMC import BatchedHMC

    dims = [8, 11, 17, 29, 45, 64, 100, 160, 256]
    targets = [TG.Funnel(d) for d in dims]
    C_, S, eps, L = 48, 10, 0.05, 10
    g = torch.Generator().manual_seed(31)
    x0 = [torch.randn(C_, d, generator=g, dtype=torch.float64) * 0.5 for d in dims]
    noise = [torch.randn(S, C_, d, generator=g, dtype=torch.float64) for d in dims]
    log_u = [torch.log(torch.rand(S, C_, generator=g, dtype=torch.float64)) for _ in dims]
    metrics = [None if i % 2 == 0 else (0.5 + torch.rand(d, generator=g, dtype=torch.float64))
               for i, d in enumerate(dims)]
    if metric == "unit":
        metrics = [None] * len(dims)
    bh = BatchedHMC(targets, C_, step_size=eps, n_leapfrog=L,
                    metrics=[m.float().cuda() if m is not None else None for m in metrics])
    assert bh.uniform_kind() == (TG.Funnel.kind if metric == "unit" else -1)
    for x, xi in zip(bh.x, x0):
        x.copy_(xi.float())
    outs = [(torch.full((C_, S + 2, d), float("nan"), device=gpu), torch.full((C_, S + 2), float("nan"), device=gpu),
             torch.full((C_, S + 2), float("nan"), device=gpu), 1) for d in dims]
    bh.run(fr.PRNGKey(0), S, outputs=outs, noise=[n.float().cuda() for n in noise],
           log_u=[u.float().cuda() for u in log_u], fused_steps=4)
    torch.cuda.synchronize()
    n_checked = 0
    for i, d in enumerate(dims):
        m = metrics[i].float().double() if metrics[i] is not None else torch.ones(d, dtype=torch.float64)
        r_pos, r_lp, r_acc, r_last, ratios = OC.take_serial_steps(
            "hmc", x0[i].float().double(), OT.funnel, noise[i].float().double(), log_u[i].float().double(), 1,
            step_size=eps, n_leapfrog=L, metric=m)
        close = ((ratios - log_u[i].float().double().T).abs() < MARGIN).any(1)
        assert int(close.sum()) <= max(2, (15 * C_) // 100), f"d={d}: {int(close.sum())} chains near the threshold"
        ok = ~close
        gp = outs[i][0][:, 1:1 + S].cpu().double()
        gl = outs[i][1][:, 1:1 + S].cpu().double()
        ga = outs[i][2][:, 1:1 + S].cpu().double()
        assert torch.equal(ga[ok], r_acc[ok]), f"d={d}: accept/reject decisions differ"
        assert ((gp[ok] - r_pos[ok]).abs() <= 2e-4 * (1 + r_pos[ok].abs())).all(), f"d={d} positions"
        assert ((gl[ok] - r_lp[ok]).abs() <= 2e-4 * (1 + r_lp[ok].abs())).all(), f"d={d} log_probs"
        assert ((bh.x[i].cpu().double()[ok] - r_last[ok]).abs() <= 2e-4 * (1 + r_last[ok].abs())).all()
        assert torch.isnan(outs[i][0][:, 0]).all() and torch.isnan(outs[i][0][:, 1 + S]).all()
        n_checked += int(ok.sum())
    assert n_checked >= 0.85 * C_ * len(dims)


def _kind_targets(kind, dims, g):
    if kind == "funnel":
        return [TG.Funnel(d) for d in dims]
    if kind == "gauss_diag":
        return [TG.GaussianDiag(torch.randn(d, generator=g), 0.5 + torch.rand(d, generator=g)) for d in dims]
    if kind == "gauss_dense":
        out = []
        for d in dims:
            a = torch.randn(d, d, generator=g, dtype=torch.float64) / d ** 0.5
            out.append(TG.GaussianDense(torch.randn(d, generator=g), a @ a.T + torch.eye(d, dtype=torch.float64)))
        return out
    if kind == "rosenbrock":
        return [TG.Rosenbrock(d) for d in dims]
    if kind == "gmm":
        return [TG.GaussianMixtureDiag(3 * torch.randn(3, d, generator=g), 0.5 + torch.rand(3, d, generator=g))
                for d in dims]
    return [TG.DualMoon(d) for d in dims]


@pytest.mark.parametrize("kind", ["funnel", "gauss_diag", "gauss_dense", "rosenbrock", "gmm", "dual_moon"])
@pytest.mark.parametrize("nv", [1, 2])
def test_batched_hmc_uniform_kernel_matches_dispatch_kernel(gpu, kind, nv):
    # A batch of one target kind with the unit metric runs the kernel
    # specialised on that kind and on each slot's lanes per chain; it must
    # reproduce the run-time dispatch kernel on the same layout: the same
    # accept decisions, and rows equal to FP32 rounding (the two builds may
    # contract a target's sums into FMAs differently: the funnel's log-density
    # differs in its last bits, an absolute 4e-6 where its terms of size ~100
    # cancel), for Philox draws and for injected noise, with lanes per chain
    # from 1 to 64 (spread 2 doubles them).
    from flowmc_amd.resource.kernel.batched_HMC import BatchedHMC

    g = torch.Generator().manual_seed(41)
    dims = [2, 5, 8, 13, 31, 64, 100, 200] if nv == 2 else [2, 3, 8, 17, 40, 64, 130, 256]
    if kind == "gauss_dense":
        dims = dims[:6]
    targets = _kind_targets(kind, dims, g)
    C_, S, eps, L = 40, 6, 0.03, 4
    x0 = [0.5 * torch.randn(C_, d, generator=g) for d in dims]
    noise = [torch.randn(S, C_, d, generator=g).cuda() for d in dims]
    log_u = [torch.log(torch.rand(S, C_, generator=g)).cuda() for _ in dims]
    runs = {}
    for spec in (True, False):
        bh = BatchedHMC(targets, C_, step_size=eps, n_leapfrog=L)
        bh.specialise = spec
        assert bh.uniform_kind() == (targets[0].kind if spec else -1)
        res = []
        for inj in (False, True):
            layout = bh.default_layout(nv, 256, 2 if inj else 1)
            for x, xi in zip(bh.x, x0):
                x.copy_(xi.cuda())
            outs = [(torch.full((C_, S, d), float("nan"), device=gpu), torch.full((C_, S), float("nan"), device=gpu),
                     torch.full((C_, S), float("nan"), device=gpu), 0) for d in dims]
            kw = dict(noise=noise, log_u=log_u) if inj else {}
            bh.run(fr.PRNGKey(5), S, outputs=outs, fused_steps=4, layout=layout, **kw)
            torch.cuda.synchronize()
            res.append(([t.clone() for o in outs for t in o[:3]], [x.clone() for x in bh.x],
                        [lp.clone() for lp in bh.logp]))
        runs[spec] = res
    for (rows_a, x_a, lp_a), (rows_b, x_b, lp_b) in zip(runs[True], runs[False]):
        for i, d in enumerate(dims):
            pa, la, aa = rows_a[3 * i:3 * i + 3]
            pb, lb, ab = rows_b[3 * i:3 * i + 3]
            assert torch.equal(aa, ab), f"d={d}: accept decisions differ"
            for a, b, tol in ((pa, pb, 1e-5), (x_a[i], x_b[i], 1e-5), (la, lb, 1e-4), (lp_a[i], lp_b[i], 1e-4)):
                assert torch.isfinite(a).all()
                assert torch.allclose(a, b, rtol=1e-5, atol=tol), (d, float((a - b).abs().max()))


def test_batched_hmc_uniform_kind_mismatch_marks_the_slots(gpu):
    # a launch that promises a uniform kind the slots do not hold (or a
    # metric on the unit-metric kernel) leaves those chains where they were
    # and sets their log-prob state to NaN instead of evaluating the wrong
    # target
    from flowmc_amd.resource.kernel.batched_HMC import BatchedHMC

    dims = [8, 40]
    bh = BatchedHMC([TG.Funnel(d) for d in dims], 64, step_size=0.05, n_leapfrog=4)
    bh.init_positions(fr.PRNGKey(0))
    x0 = [x.clone() for x in bh.x]
    bh.uniform_kind = lambda: TG.GaussianDiag.kind  # the wrong promise
    bh.run(fr.PRNGKey(1), 4, fused_steps=4)
    torch.cuda.synchronize()
    for x, xi, lp in zip(bh.x, x0, bh.logp):
        assert torch.equal(x, xi)
        assert bool(torch.isnan(lp).all())


def test_batched_hmc_keeps_the_funnel_law_at_full_size(gpu):
    # 32 funnels (d 8..256, log-uniform), 256 chains each, L = 10, eps = 0.05,
    # chains started from exact funnel draws: after 96 HMC steps the pooled
    # final states must still have v ~ N(0, 9) and x_i e^{-v/2} ~ N(0, 1)
    # (an exact sampler leaves its target invariant). Tolerances: 5 standard
    # errors of the 8192 independent chains.
    from flowmc_amd.resource.kernel.batched_HMC import BatchedHMC

    dims = TG.funnel_dims(0, n_targets=32, lo=8, hi=256)
    C_ = 256
    bh = BatchedHMC([TG.Funnel(d) for d in dims], C_, step_size=0.05, n_leapfrog=10)
    g = torch.Generator().manual_seed(41)
    for x in bh.x:
        v = 3.0 * torch.randn(C_, generator=g)
        rest = torch.randn(C_, x.shape[1] - 1, generator=g) * torch.exp(0.5 * v)[:, None]
        x.copy_(torch.cat([v[:, None], rest], 1))
    bh.run(fr.PRNGKey(5), 96, fused_steps=16)
    torch.cuda.synchronize()
    v = torch.cat([x[:, 0].double().cpu() for x in bh.x])
    n = v.numel()
    assert n == 32 * C_
    assert abs(float(v.mean())) < 5 * 3 / n ** 0.5
    assert abs(float(v.var()) - 9) < 5 * 9 * (2 / n) ** 0.5
    u2 = torch.cat([(x[:, 1:].double().cpu() ** 2 * torch.exp(-x[:, :1].double().cpu())).mean(1) for x in bh.x])
    # E[x_i^2 e^{-v}] = 1 for each chain's average over its d - 1 conditionals
    assert abs(float(u2.mean()) - 1) < 5 * float(u2.std()) / n ** 0.5
    lp = torch.cat([lp.double().cpu() for lp in bh.logp])
    assert bool(torch.isfinite(lp).all())


@pytest.mark.parametrize("name,d", [("gauss_dense", 5), ("gmm", 16), ("funnel", 40), ("rosenbrock", 9)])
def test_hmc_dense_condition_matrix_matches_oracle(gpu, name, d):
    """Dense condition matrix M (HMC.py:118-136): momentum xi @ chol(inv(M)).T,
    broadcast kinetic 1/2 sum_j colsum_j(M) p_j^2, fused kernel against the
    FP64 oracle on injected noise."""
    dt = {"gauss_dense": 0.2, "gmm": 0.1, "funnel": 0.03, "rosenbrock": 0.02}[name]
    acc = _compare_traj("hmc", name, d, 96, 16, dt, n_leapfrog=6, dense=True)
    assert float(acc.mean()) > 0.05


def test_hmc_dense_condition_matrix_generic_target_matches_oracle(gpu):
    """The per-step path (a user target differentiated by torch) with a dense
    condition matrix: fmc_hmc_momentum's staged-LDS product against the oracle."""
    d = 7
    g = torch.Generator().manual_seed(5)
    mean = torch.randn(d, generator=g, dtype=torch.float64)
    prec = 0.5 + torch.rand(d, generator=g, dtype=torch.float64)

    def user_logpdf(x, data):
        return -0.5 * torch.sum(data["prec"] * (x - data["mean"]) ** 2)

    fn = lambda z: OT.gauss_diag(z, mean, prec)
    data = {"mean": mean.float().cuda(), "prec": prec.float().cuda()}
    C_, S, dt, L = 64, 10, 0.25, 5
    x0 = torch.randn(C_, d, generator=g, dtype=torch.float64) * 0.5
    noise = torch.randn(S, C_, d, generator=g, dtype=torch.float64)
    log_u = torch.log(torch.rand(S, C_, generator=g, dtype=torch.float64))
    M = dense_condition_matrix(d, g).float().double()
    pos = torch.zeros(C_, S, d, device=gpu)
    lps = torch.zeros(C_, S, device=gpu)
    accs = torch.zeros(C_, S, device=gpu)
    HMC(M.float(), dt, L).sample_steps(fr.PRNGKey(0), x0.float().cuda(), LogPDF(user_logpdf, n_dims=d), data, S, 1,
                                      StepOutputs(pos, lps, accs, 0), noise=noise.float().cuda(),
                                      log_u=log_u.float().cuda())
    r_pos, r_lp, r_acc, _, ratios = OC.take_serial_steps("hmc", x0.float().double(), fn, noise.float().double(),
                                                         log_u.float().double(), 1, step_size=dt, n_leapfrog=L,
                                                         metric=M)
    ok = ~((ratios - log_u.float().double().T).abs() < MARGIN).any(1)
    assert int((~ok).sum()) <= 4
    assert torch.equal(accs.cpu().double()[ok], r_acc[ok])
    assert ((pos.cpu().double()[ok] - r_pos[ok]).abs() <= 2e-4 * (1 + r_pos[ok].abs())).all()
    assert 0.05 < float(r_acc.mean())


def test_batched_hmc_dense_condition_matrix_matches_oracle(gpu):
    """A batch mixing unit, diagonal and dense condition matrices."""
    from flowmc_amd.resource.kernel.batched_HMC import BatchedHMC

    dims = [9, 24, 40]
    g = torch.Generator().manual_seed(12)
    metrics = [None, 0.5 + torch.rand(24, generator=g, dtype=torch.float64), dense_condition_matrix(40, g)]
    C_, S, eps, L = 48, 8, 0.04, 10
    x0 = [torch.randn(C_, d, generator=g, dtype=torch.float64) * 0.5 for d in dims]
    noise = [torch.randn(S, C_, d, generator=g, dtype=torch.float64) for d in dims]
    log_u = [torch.log(torch.rand(S, C_, generator=g, dtype=torch.float64)) for _ in dims]
    bh = BatchedHMC([TG.Funnel(d) for d in dims], C_, step_size=eps, n_leapfrog=L,
                    metrics=[None, metrics[1].float().cuda(), metrics[2].float()])
    for x, xi in zip(bh.x, x0):
        x.copy_(xi.float())
    outs = [(torch.zeros(C_, S, d, device=gpu), torch.zeros(C_, S, device=gpu), torch.zeros(C_, S, device=gpu), 0)
            for d in dims]
    bh.run(fr.PRNGKey(0), S, outputs=outs, noise=[n.float().cuda() for n in noise],
           log_u=[u.float().cuda() for u in log_u], fused_steps=4)
    torch.cuda.synchronize()
    for i, d in enumerate(dims):
        m = torch.ones(d, dtype=torch.float64) if metrics[i] is None else metrics[i].float().double()
        r_pos, _, r_acc, _, ratios = OC.take_serial_steps("hmc", x0[i].float().double(), OT.funnel,
                                                          noise[i].float().double(), log_u[i].float().double(), 1,
                                                          step_size=eps, n_leapfrog=L, metric=m)
        ok = ~((ratios - log_u[i].float().double().T).abs() < MARGIN).any(1)
        assert int((~ok).sum()) <= 4
        ga = outs[i][2].cpu().double()
        assert torch.equal(ga[ok], r_acc[ok]), f"d={d}"
        gp = outs[i][0].cpu().double()
        assert ((gp[ok] - r_pos[ok]).abs() <= 2e-4 * (1 + r_pos[ok].abs())).all(), f"d={d}"
