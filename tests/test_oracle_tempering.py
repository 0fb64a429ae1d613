"""Pins of the parallel-tempering oracle (oracle/tempering.py) and the host
semantics of TemperedPDF; CPU only.

The exchange sweep is checked by what the mathematics fixes: each pairwise
Metropolis swap leaves the product of the tempered densities invariant, so
walkers drawn exactly from N(0, T_t) in slot t keep variance T_t after a sweep
(a wrong sign or a swapped temperature index sends the wide walkers to the
cold slot); plus known answers that pin the order of the sweep.
"""

import math

import pytest
import torch

from flowmc_amd.resource.logPDF import TemperedPDF
from oracle import chains as OC
from oracle import targets as OT
from oracle import tempering as OPT


def std_normal(x):
    return -0.5 * (x * x).sum(-1)


def test_exchange_leaves_tempered_product_invariant():
    g = torch.Generator().manual_seed(0)
    temps = torch.tensor([1.0, 1.7, 3.0, 5.0], dtype=torch.float64)
    C, T, d = 200_000, 4, 3
    x = torch.randn(C, T, d, generator=g, dtype=torch.float64) * temps.sqrt()[None, :, None]
    L = std_normal(x)
    log_u = torch.log(torch.rand(C, T - 1, generator=g, dtype=torch.float64))
    y, Ly, acc = OPT.exchange(x, L, temps, log_u)
    var = y.pow(2).mean((0, 2))
    assert torch.allclose(var, temps, rtol=1.5e-2), var
    assert torch.allclose(Ly, std_normal(y))
    # swaps do happen (otherwise the invariance check is vacuous)
    assert 0.2 < float(acc.double().mean()) < 0.95
    # a plausible mistake breaks it: comparing with the opposite sign
    y_bad, _, _ = OPT.exchange(x, -L, temps, log_u)
    var_bad = y_bad.pow(2).mean((0, 2))
    assert float(var_bad[0]) > 1.2


def test_cold_slot_takes_the_better_walker():
    temps = torch.tensor([1.0, 2.0], dtype=torch.float64)
    x = torch.tensor([[[0.0], [1.0]]], dtype=torch.float64)
    # L_hot > L_cold: ratio = (1/2 - 1)(-5 - 0) = 2.5 > 0 -> accepted for any u < 1
    y, L, acc = OPT.exchange(x, torch.tensor([[-5.0, 0.0]], dtype=torch.float64), temps,
                             torch.tensor([[math.log(0.999)]], dtype=torch.float64))
    assert bool(acc[0, 0]) and float(y[0, 0, 0]) == 1.0 and L.tolist() == [[0.0, -5.0]]
    # L_cold > L_hot: ratio = -2.5 < log(0.5)
    y, L, acc = OPT.exchange(x, torch.tensor([[0.0, -5.0]], dtype=torch.float64), temps,
                             torch.tensor([[math.log(0.5)]], dtype=torch.float64))
    assert not bool(acc[0, 0]) and torch.equal(y, x)


def test_walker_climbs_several_slots_in_one_sweep():
    temps = torch.tensor([1.0, 2.0, 4.0], dtype=torch.float64)
    x = torch.tensor([[[10.0], [20.0], [30.0]]], dtype=torch.float64)
    L = torch.tensor([[-10.0, 0.0, 0.0]], dtype=torch.float64)
    # pair 0: (1/2 - 1)(-10 - 0) = 5; pair 1 sees the moved walker: (1/4 - 1/2)(-10 - 0) = 2.5
    y, Ly, acc = OPT.exchange(x, L, temps, torch.full((1, 2), -1e-3, dtype=torch.float64))
    assert acc.tolist() == [[True, True]]
    assert y[0, :, 0].tolist() == [20.0, 30.0, 10.0]
    assert Ly.tolist() == [[0.0, 0.0, -10.0]]


def test_equal_temperatures_accept_every_swap():
    g = torch.Generator().manual_seed(1)
    x = torch.randn(5, 4, 2, generator=g, dtype=torch.float64)
    L = torch.randn(5, 4, generator=g, dtype=torch.float64)
    log_u = torch.log(torch.rand(5, 3, generator=g, dtype=torch.float64))
    y, Ly, acc = OPT.exchange(x, L, torch.full((4,), 2.0, dtype=torch.float64), log_u)
    assert bool(acc.all())
    # every pair swaps: slot 0's walker rides to the top, the others shift down
    assert torch.equal(y, torch.cat([x[:, 1:], x[:, :1]], 1))
    assert torch.equal(Ly, torch.cat([L[:, 1:], L[:, :1]], 1))


def test_adapt_temperature_known_answers():
    temps = torch.tensor([1.0, 2.0, 3.0, 4.0], dtype=torch.float64)
    # equal acceptance on every pair: no damping, ladder unchanged
    assert torch.equal(OPT.adapt_temperature(temps, torch.ones(50, 3)), temps)
    # hand-worked: C = 100, rates (1, 0, 1/2) -> damping (1, -1/2);
    # T1 = 1 + (2 - 1) e, T2 = T1 + (3 - 2) e^-1/2, T3 = 4 unchanged
    acc = torch.zeros(100, 3)
    acc[:, 0] = 1
    acc[:50, 2] = 1
    new = OPT.adapt_temperature(temps, acc)
    t1 = 1 + math.e
    assert new.tolist() == pytest.approx([1.0, t1, t1 + math.exp(-0.5), 4.0], rel=1e-12)
    # the reference's own call shape: decisions [n_chains, n_temps]
    t = torch.arange(5, dtype=torch.float64) * 0.3 + 1
    assert torch.equal(OPT.adapt_temperature(t, torch.ones(7, 5)), t)


def test_tempered_rows_scale_by_inverse_temperature():
    f = OPT.tempered_rows(std_normal, [1.0, 4.0], 3)
    x = torch.ones(6, 2, dtype=torch.float64)
    assert f(x).tolist() == [-1.0, -0.25] * 3


def test_tempered_mala_targets_the_tempered_density():
    # walkers started from the exact tempered density stay there:
    # slot t of N(0, 1) tempered at T_t is N(0, T_t)
    g = torch.Generator().manual_seed(2)
    temps = torch.tensor([1.0, 4.0], dtype=torch.float64)
    C, d, S = 4000, 2, 40
    x0 = torch.randn(C, 2, d, generator=g, dtype=torch.float64) * temps.sqrt()[None, :, None]
    noise = torch.randn(S, 2 * C, d, generator=g, dtype=torch.float64)
    log_u = torch.log(torch.rand(S, 2 * C, generator=g, dtype=torch.float64))
    x, lp, acc, _ = OPT.ensemble_steps("mala", x0, std_normal, temps, noise, log_u, step_size=0.9)
    var = x.pow(2).mean((0, 2))
    assert torch.allclose(var, temps, rtol=6e-2), var
    assert torch.allclose(lp, std_normal(x) / temps[None])
    assert 0.3 < float(acc.double().mean()) < 1.0


def test_unit_temperature_equals_untempered_steps():
    g = torch.Generator().manual_seed(3)
    C, d, S = 6, 4, 5
    x0 = torch.randn(C, 1, d, generator=g, dtype=torch.float64)
    noise = torch.randn(S, C, d, generator=g, dtype=torch.float64)
    log_u = torch.log(torch.rand(S, C, generator=g, dtype=torch.float64))
    fn = OT.rosenbrock
    x, lp, acc, _ = OPT.ensemble_steps("grw", x0, fn, [1.0], noise, log_u, step_size=0.05)
    _, lps, accs, last, _ = OC.take_serial_steps("grw", x0[:, 0], fn, noise, log_u, step_size=0.05)
    assert torch.equal(x[:, 0], last) and torch.equal(lp[:, 0], lps[:, -1]) and torch.equal(acc[:, 0], accs)


# ---------------------------------------------------------- host semantics --
def posterior(x, data):
    return -torch.sum(torch.square(x - data["data"]))


def test_tempered_pdf_matches_reference_semantics():
    # test_resources.py:59-81: at x = data the likelihood is 0, so every tempered value is 0
    logpdf = TemperedPDF(posterior, lambda x, data: torch.zeros(1), n_dims=3, n_temps=5, max_temp=100)
    inputs = torch.ones(5, 3)
    data = {"data": torch.ones(3)}
    temps = torch.arange(5) + 1.0
    vals = logpdf.tempered_log_pdf(temps, inputs, data)
    assert vals.shape == (5,) and torch.equal(vals, logpdf(inputs, data))
    # away from the mode: L / T + prior, and calling the object gives L (no prior, untempered)
    x = torch.tensor([[0.0, 1.0, 3.0]] * 5)
    L = posterior(x[0], data)
    prior = TemperedPDF(posterior, lambda x, data: x.sum() * 0 + 2.0, n_dims=3)
    got = prior.tempered_log_pdf(temps, x, data)
    assert torch.allclose(got, L / temps + 2.0)
    assert torch.allclose(prior(x, data), L.expand(5))
    assert not prior.fused_tempering and not logpdf.fused_tempering
    assert repr(prior) == "TemperedPDF with 3 dimensions"
