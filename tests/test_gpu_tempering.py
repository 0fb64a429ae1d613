"""GPU parity of parallel tempering against the FP64 oracle (oracle/tempering.py).

Tolerances as for the chain kernels (FP32 kernels vs FP64 oracle on identical
injected noise): positions within 2e-4 * (1 + |x|), log-probs within
2e-4 * (1 + |lp|), step / swap decisions identical except on chains where
some oracle log-ratio lies within MARGIN of its log-uniform (counted, rare).
The exchange gathers rows, so positions of agreeing chains match bit for bit.
"""

import math

import pytest
import torch

import flowmc_amd._native as N
from flowmc_amd import random as fr
from flowmc_amd import targets as TG
from flowmc_amd.resource.buffers import Buffer
from flowmc_amd.resource.kernel.Gaussian_random_walk import GaussianRandomWalk
from flowmc_amd.resource.kernel.HMC import HMC
from flowmc_amd.resource.kernel.MALA import MALA
from flowmc_amd.resource.logPDF import LogPDF, TemperedPDF
from flowmc_amd.resource.states import State
from flowmc_amd.resource_strategy_bundle.RQSpline_MALA_PT import RQSpline_MALA_PT_Bundle
from flowmc_amd.Sampler import Sampler
from flowmc_amd.strategy.parallel_tempering import ParallelTempering
from oracle import philox as OP
from oracle import tempering as OPT
from test_gpu_chains import make_target

pytestmark = pytest.mark.gpu

MARGIN = 2e-4


def _proposal(kind, dt, d, metric=None):
    if kind == "mala":
        return MALA(dt, fused_steps=16)
    if kind == "grw":
        return GaussianRandomWalk(dt, fused_steps=16)
    return HMC(metric.to(torch.float32), dt, 4, fused_steps=8)


def _tempered_vs_oracle(kind, name, d, C_, T, S, dt, generic=False, seed=0):
    tgt, fn = make_target(name, d, seed=seed)
    g = torch.Generator().manual_seed(seed + 5)
    temps = torch.linspace(1.0, 4.0, T, dtype=torch.float64).float()
    x0 = (torch.randn(C_, T, d, generator=g, dtype=torch.float64) * 0.5).float()
    noise = torch.randn(S, C_ * T, d, generator=g, dtype=torch.float64).float()
    log_u = torch.log(torch.rand(S, C_ * T, generator=g, dtype=torch.float64)).float()
    metric = (0.5 + torch.rand(d, generator=g, dtype=torch.float64)).float() if kind == "hmc" else None
    prop = _proposal(kind, dt, d, metric)
    if generic:
        # the same target as a torch callable: per-step kernels around autograd
        def lik(x, data):
            return tgt(x[None])[0] if x.ndim == 1 else tgt(x)
        tpdf = TemperedPDF(LogPDF(lik, n_dims=d, batched=True), None, n_dims=d)
        assert not tpdf.fused_tempering
    else:
        tpdf = TemperedPDF(LogPDF(tgt, n_dims=d), None, n_dims=d)
        assert tpdf.fused_tempering
    acc = torch.full((C_ * T, S + 2), float("nan"), device="cuda")
    pos, lp = prop.sample_tempered(fr.PRNGKey(0), x0.cuda(), tpdf, temps.cuda(), None, S, noise=noise.cuda(),
                                   log_u=log_u.cuda(), accepts=acc)
    torch.cuda.synchronize()
    kw = dict(step_size=dt)
    if kind == "hmc":
        kw.update(n_leapfrog=4, metric=metric.double())
    r_x, r_lp, r_acc, ratios = OPT.ensemble_steps(kind, x0.double(), fn, temps.double(), noise.double(),
                                                  log_u.double(), **kw)
    close = ((ratios - log_u.double().T).abs() < MARGIN).any(1).reshape(C_, T)
    assert int(close.sum()) <= max(2, (15 * C_ * T) // 100)
    ok = ~close
    ga = acc[:, :S].cpu().reshape(C_, T, S)
    assert torch.equal(ga[ok], r_acc[ok].double().float()), "step decisions differ"
    assert torch.isnan(acc[:, S:]).all()
    gx, gl = pos.cpu().double(), lp.cpu().double()
    assert ((gx[ok] - r_x[ok]).abs() <= 2e-4 * (1 + r_x[ok].abs())).all()
    assert ((gl[ok] - r_lp[ok]).abs() <= 2e-4 * (1 + r_lp[ok].abs())).all()
    return ga


@pytest.mark.parametrize("kind,name,d,dt", [("mala", "gmm", 16, 0.3), ("mala", "rosenbrock", 12, 0.02),
                                            ("grw", "gauss_dense", 5, 0.4), ("hmc", "funnel", 10, 0.1)])
def test_fused_tempered_steps_match_oracle(gpu, kind, name, d, dt):
    acc = _tempered_vs_oracle(kind, name, d, 96, 4, 40, dt)
    assert 0.05 < float(acc.mean()) < 1.0


@pytest.mark.parametrize("kind", ["mala", "grw", "hmc"])
def test_generic_tempered_steps_match_oracle(gpu, kind):
    _tempered_vs_oracle(kind, "gauss_diag", 6, 32, 3, 12, {"mala": 0.4, "grw": 0.5, "hmc": 0.2}[kind],
                        generic=True)


def test_tempered_philox_mode_generic_equals_fused(gpu):
    # no injected noise: the per-step kernels and the fused launch draw the same Philox streams
    d, C_, T, S = 5, 40, 3, 20
    mean = torch.arange(d, dtype=torch.float32)
    native = TemperedPDF(LogPDF(TG.GaussianDiag(mean, None), n_dims=d), None)

    def lik(x, data):
        return -0.5 * torch.sum((x - data["data"]) ** 2)

    generic = TemperedPDF(LogPDF(lik, n_dims=d), None)
    temps = torch.tensor([1.0, 2.0, 5.0], device="cuda")
    x0 = torch.randn(C_, T, d, device="cuda")
    out = []
    for tp in (native, generic):
        acc = torch.zeros(C_ * T, S, device="cuda")
        pos, lp = MALA(0.5).sample_tempered(fr.PRNGKey(9), x0, tp, temps, {"data": mean.cuda()}, S, accepts=acc)
        out.append((pos, lp, acc))
    (p1, l1, a1), (p2, l2, a2) = out
    agree = (a1 == a2).all(1).reshape(C_, T)
    assert float(agree.float().mean()) > 0.95
    assert torch.allclose(p1[agree], p2[agree], atol=1e-4, rtol=1e-4)
    assert torch.allclose(l1[agree], l2[agree], atol=1e-4, rtol=1e-4)


def _exchange(x, L, temps, log_u=None, key=(0, 0)):
    C_, T, d = x.shape
    xi = x.contiguous().cuda()
    out = torch.full_like(xi, float("nan"))
    Lg = L.contiguous().cuda().clone()
    acc = torch.full((C_, max(T - 1, 1)), float("nan"), device="cuda")
    tg = temps.contiguous().cuda()
    lu = log_u.contiguous().cuda() if log_u is not None else None
    N.call("fmc_pt_exchange", N.stream_ptr(), xi.data_ptr(), out.data_ptr(), Lg.data_ptr(), tg.data_ptr(), C_, T, d,
           key[0], key[1], lu.data_ptr() if lu is not None else None, acc.data_ptr())
    torch.cuda.synchronize()
    return out.cpu(), Lg.cpu(), acc[:, :T - 1].cpu()


@pytest.mark.parametrize("C_,T,d", [(1, 2, 1), (300, 5, 7), (129, 16, 64), (70, 64, 3), (5, 1, 4)])
def test_exchange_matches_oracle(gpu, C_, T, d):
    g = torch.Generator().manual_seed(T * 100 + d)
    temps = torch.linspace(1.0, 6.0, T).float() if T > 1 else torch.ones(1)
    x = torch.randn(C_, T, d, generator=g)
    L = (torch.randn(C_, T, generator=g) * 3).float()
    lu = torch.log(torch.rand(C_, max(T - 1, 0), generator=g))
    y, Ly, acc = _exchange(x, L, temps, lu if T > 1 else None)
    r_y, r_L, r_acc = OPT.exchange(x.double(), L.double(), temps.double(), lu.double())
    # margin: a decision whose ratio is within rounding of log u
    Tn = temps.double()
    if T > 1:
        ok = torch.ones(C_, dtype=torch.bool)
        Lc = L.double().clone()
        for i in range(T - 1):  # the oracle's ratios, replayed
            ratio = (1 / Tn[i + 1] - 1 / Tn[i]) * (Lc[:, i] - Lc[:, i + 1])
            ok &= (ratio - lu[:, i].double()).abs() > 1e-5
            a = r_acc[:, i]
            tmp = Lc[a, i].clone()
            Lc[a, i] = Lc[a, i + 1]
            Lc[a, i + 1] = tmp
        assert int((~ok).sum()) <= 2
        assert torch.equal(acc[ok].bool(), r_acc[ok])
    else:
        ok = torch.ones(C_, dtype=torch.bool)
    assert torch.equal(y[ok], r_y[ok].float())
    assert torch.equal(Ly[ok], r_L[ok].float())


def test_exchange_philox_mode_equals_injected(gpu):
    C_, T, d = 200, 6, 5
    key = (77, 1234)
    g = torch.Generator().manual_seed(8)
    temps = torch.linspace(1.0, 3.0, T)
    x = torch.randn(C_, T, d, generator=g)
    L = torch.randn(C_, T, generator=g) * 2
    lu = torch.empty(T - 1, C_, device="cuda")
    N.call("fmc_rng_log_uniform", N.stream_ptr(), key[0], key[1], C_, T - 1, 0, OP.RNG_PT_SWAP, lu.data_ptr())
    a = _exchange(x, L, temps, None, key)
    b = _exchange(x, L, temps, lu.T.contiguous().cpu(), key)
    for u, v in zip(a, b):
        assert torch.equal(u, v)
    # and the stream is Philox stream 9 with counter (chain, pair)
    want = OP.log_uniforms(key, C_, T - 1, 0, OP.RNG_PT_SWAP)
    assert torch.allclose(lu.cpu().double(), torch.as_tensor(want), atol=1e-6)


@pytest.mark.parametrize("C_,T", [(7, 5), (1000, 8), (4096, 3), (3, 2), (300, 64)])
def test_adapt_matches_oracle(gpu, C_, T):
    g = torch.Generator().manual_seed(C_ + T)
    temps = torch.sort(1 + 4 * torch.rand(T, generator=g)).values
    temps[0] = 1.0
    acc = (torch.rand(C_, T - 1, generator=g) < torch.rand(T - 1, generator=g)).float()
    tg = temps.cuda().contiguous()
    ag = acc.cuda().contiguous()
    N.call("fmc_pt_adapt", N.stream_ptr(), tg.data_ptr(), ag.data_ptr(), C_, T)
    want = OPT.adapt_temperature(temps.double(), acc.double())
    # exp of (100/C) * rate differences: the damping can be large for small C, so relative tolerance
    assert torch.allclose(tg.cpu().double(), want, rtol=2e-5, atol=0)
    assert float(tg[0]) == float(temps[0]) and float(tg[-1]) == float(temps[-1])


# ------------------------------------------------------------- strategy ----
def log_posterior(x, data={}):
    return -0.5 * torch.sum((x - data["data"]) ** 2)


def _pt_resources(training, n_temps=5, n_dims=3, n_chains=7):
    # test_strategies.py:337-390
    mala = MALA(1.0)
    logpdf = TemperedPDF(log_posterior, lambda x, data: torch.tensor(0.0), n_dims=n_dims, n_temps=n_temps)
    k1, k2 = fr.split(fr.PRNGKey(42))
    initial_position = fr.normal(k1, (n_chains, n_dims))
    tempered_positions = Buffer("tempered_positions", (n_chains, n_temps - 1, n_dims), 2)
    tempered_positions.update_buffer(fr.normal(k2, (n_chains, n_temps - 1, n_dims)))
    temperatures = Buffer("temperatures", (n_temps,), 0)
    temperatures.update_buffer(torch.arange(n_temps) + 1.0)
    state = State({"target_positions": "tempered_positions", "target_log_prob": "logpdf",
                   "target_temperatures": "temperatures", "training": training}, name="sampler_state")
    resources = {"logpdf": logpdf, "MALA": mala, "tempered_positions": tempered_positions,
                 "temperatures": temperatures, "sampler_state": state}
    strat = ParallelTempering(n_steps=4, tempered_logpdf_name="logpdf", kernel_name="MALA",
                              tempered_buffer_names=["tempered_positions", "temperatures"],
                              state_name="sampler_state")
    return resources, strat, initial_position


def test_parallel_tempering_strategy_training_and_production(gpu):
    data = {"data": torch.arange(3, dtype=torch.float32, device="cuda")}
    for training in (False, True):
        resources, strat, x0 = _pt_resources(training)
        tp0 = resources["tempered_positions"].data.clone()
        t0 = resources["temperatures"].data.clone()
        key, res, pos = strat(fr.PRNGKey(1), resources, x0, data)
        assert pos.shape == (7, 3) and torch.isfinite(pos).all()
        if training:
            assert not torch.equal(res["tempered_positions"].data, tp0)
            assert res["temperatures"].data.shape == (5,)
            assert float(res["temperatures"].data[0]) == 1.0 and float(res["temperatures"].data[-1]) == 5.0
        else:  # production leaves the tempered walkers and the ladder alone
            assert torch.equal(res["tempered_positions"].data, tp0)
            assert torch.equal(res["temperatures"].data, t0)


def test_parallel_tempering_helpers_shapes(gpu):
    resources, strat, x0 = _pt_resources(True)
    data = {"data": torch.arange(3, dtype=torch.float32, device="cuda")}
    positions = torch.cat([x0.cuda()[:, None], resources["tempered_positions"].data], 1)
    temps = torch.arange(5, device="cuda") * 0.3 + 1
    p, lp, acc = strat._ensemble_step(resources["MALA"], fr.PRNGKey(3), positions, resources["logpdf"], temps, data)
    assert p.shape == (7, 5, 3) and lp.shape == (7, 5) and acc.shape == (7, 5, 4)
    want = resources["logpdf"].tempered_log_pdf(temps[None].expand(7, 5), p, data)
    assert torch.allclose(lp, want, atol=1e-5, rtol=1e-5)
    p2, L2, sw = strat._exchange(fr.PRNGKey(4), p, resources["logpdf"], temps, data)
    assert sw.shape == (7, 4) and torch.allclose(L2, resources["logpdf"](p2, data), atol=1e-5)
    # the walkers are a permutation of the input walkers within each chain
    assert torch.allclose(p2.sort(1).values.sum(), p.sort(1).values.sum())
    new = strat._adapt_temperature(temps, torch.ones(7, 5))
    assert torch.allclose(new, temps)


def test_per_walker_helpers_follow_the_reference_tests(gpu):
    # test_strategies.py:395-470: one walker's step body, its n_steps scan, one
    # chain's ensemble step, and the exchange / adaptation of one chain
    resources, strat, x0 = _pt_resources(False)
    mala, logpdf = resources["MALA"], resources["logpdf"]
    data = {"data": torch.arange(3, dtype=torch.float32, device="cuda")}
    key = fr.PRNGKey(42)
    position = x0[0]
    log_prob = logpdf(position, data)
    carry, extras = strat._individual_step_body(mala, (key, position, log_prob, logpdf, torch.tensor(1.0), data), None)
    assert carry[1].shape == (3,) and extras[0].shape == (3,)
    assert torch.allclose(carry[2], logpdf.tempered_log_pdf(1.0, carry[1], data), atol=1e-5)
    walkers = torch.cat([x0.cuda()[:, None], resources["tempered_positions"].data], 1)
    pos, lp, acc = strat._individal_step(mala, key, walkers[0, 0], logpdf, torch.tensor(1.0), data)
    assert pos.shape == (3,) and acc.shape == (4,)
    temps = torch.arange(5, dtype=torch.float32) + 1.0
    p1, l1, a1 = strat._ensemble_step(mala, key, walkers[0], logpdf, temps, data)
    assert p1.shape == (5, 3) and l1.shape == (5,) and a1.shape == (5, 4)
    # the one-chain call is the batched call restricted to that chain
    pb, lb, ab = strat._ensemble_step(mala, key, walkers[:1], logpdf, temps, data)
    assert torch.equal(p1, pb[0]) and torch.equal(a1, ab[0])
    y, L, sw = strat._exchange(fr.PRNGKey(7), p1, logpdf, torch.arange(5) * 0.3 + 1, data)
    assert y.shape == (5, 3) and L.shape == (5,) and sw.shape == (4,)
    assert torch.allclose(y.sort(0).values.sum(), p1.sort(0).values.sum())
    new = strat._adapt_temperature(torch.arange(5) * 0.3 + 1, torch.ones(7, 5))
    assert new.shape == (5,)


def test_tempered_sampler_recovers_bimodal_weights(gpu):
    # two well-separated modes: the cold chains alone stay stuck, tempering lets them hop
    d, C_, T = 2, 512, 6
    t = TG.GaussianMixtureDiag(torch.tensor([[-4.0, 0.0], [4.0, 0.0]]), torch.full((2, d), 0.7),
                               torch.tensor([0.5, 0.5]))
    tpdf = TemperedPDF(LogPDF(t, n_dims=d), None)
    temps = Buffer("temperatures", (T,), 0)
    temps.update_buffer(torch.tensor([1.0, 2.0, 4.0, 8.0, 16.0, 32.0]))
    tpos = Buffer("tempered_positions", (C_, T - 1, d), 2)
    x = torch.tensor([-4.0, 0.0], device="cuda").expand(C_, d).contiguous()  # every chain starts in one mode
    tpos.update_buffer(x[:, None].expand(C_, T - 1, d))
    state = State({"training": True}, name="s")  # training: the tempered walkers carry over between calls
    res = {"lp": tpdf, "K": MALA(0.8), "tp": tpos, "T": temps, "s": state}
    strat = ParallelTempering(50, "lp", "K", ["tp", "T"], "s")
    key = fr.PRNGKey(0)
    right = []
    for it in range(100):
        key, res, x = strat(key, res, x, {})
        if it >= 40:
            right.append((x[:, 0] > 0).float().mean())
    frac = float(torch.stack(right).mean())
    assert abs(frac - 0.5) < 0.1, frac


def test_pt_bundle_structure_and_run(gpu, capsys):
    n_chains, n_dims = 64, 3
    t = TG.GaussianDiag(torch.zeros(n_dims), None)
    b = RQSpline_MALA_PT_Bundle(fr.PRNGKey(0), n_chains, n_dims, LogPDF(t, n_dims=n_dims), 10, 5, 2, 1, 3,
                                rq_spline_hidden_units=[32, 32], n_max_examples=500, batch_size=200)
    assert "n_tempered_steps value is not valid. Setting to n_local_steps" in capsys.readouterr().out
    assert repr(b) == "RQSpline MALA PT Bundle"
    assert b.strategy_order[0] == "initialize_tempered_positions"
    assert b.strategy_order[1:8] == ["parallel_tempering", "local_stepper", "update_global_step", "model_trainer",
                                     "update_model", "global_stepper", "update_local_step"]
    assert b.strategy_order[-5:] == ["parallel_tempering", "local_stepper", "update_global_step", "global_stepper",
                                     "update_local_step"]
    assert len(b.strategy_order) == 1 + 2 * 7 + 2 + 5
    assert b.resources["temperatures"].data.tolist() == [1.0, 2.0, 3.0, 4.0, 5.0]
    assert b.resources["temperatures"].name == "temperature"
    assert b.resources["tempered_positions"].data.shape == (n_chains, 4, n_dims)
    assert b.strategies["parallel_tempering"].n_steps == 10
    assert b.resources["tempered_logpdf"].fused_tempering
    s = Sampler(n_dims, n_chains, fr.PRNGKey(1), resource_strategy_bundles=b)
    last = s.sample(fr.normal(fr.PRNGKey(2), (n_chains, n_dims)), {})
    assert last.shape == (n_chains, n_dims) and torch.isfinite(last).all()
    prod = s.resources["positions_production"].data
    assert torch.isfinite(prod).all()
    temps = s.resources["temperatures"].data
    assert float(temps[0]) == 1.0 and float(temps[-1]) == 5.0 and torch.isfinite(temps).all()
    assert math.isfinite(float(s.resources["tempered_positions"].data.sum()))
