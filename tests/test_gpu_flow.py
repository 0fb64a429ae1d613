"""GPU parity of the RQ-spline flow kernels against the FP64 oracle.

Tolerances (stated per precision of the conditioner GEMMs; errors scaled as
|a - b| / (1 + |b|)):
  FP32 conditioner: log_prob / log-det / positions within 2e-4, every
                    element, at every depth (the forward-then-exact-inverse
                    round trip, a property of the flow's conditioning rather
                    than a comparison with the oracle, within 5e-3: the inverse
                    divides FP32 rounding by the local slope); parameter gradients within
                    2e-3 * max|grad| (FP32 accumulation over the batch vs FP64).
  BF16 conditioner, against the oracle that rounds GEMM operands to BF16 as
                    the kernels do: 99 % of elements within 2e-3 (up to 4
                    layers) or 1e-2 (8-10 layers: one rounding tie per layer
                    adds up along the depth), and every element within 3e-2
                    at full depth and 0.15 for the shallow cases' wilder
                    parameters (output logits of unit variance: a rounding tie
                    in a hidden activation moves a steep or nearly flat spline
                    segment, and the inverse divides by the local slope);
                    gradients: relative L2 error <= 1e-2 against that oracle,
                    and cosine similarity >= 0.999 with relative L2 error
                    <= 5e-2 against the exact FP64 gradient.
Full-depth cases (8 and 10 layers) and the cases with one, three or four
hidden layers draw moderate per-layer parameters (ScalarAffine within +-5 %,
output weights N(0, 0.1 / fan_in)): with the two-hidden-layer shallow cases'
per-layer magnitudes, ten layers compound into a flow whose FP32 forward
overflows on 3-sigma inputs, and single draws give nearly flat spline
segments whose inverse turns a last-bit difference into an O(1) one.
"""

import ctypes as C

import pytest
import torch

import flowmc_amd._native as N
from flowmc_amd import random as fr
from flowmc_amd.resource.buffers import Buffer
from flowmc_amd.resource.kernel.NF_proposal import NFProposal
from flowmc_amd.resource.logPDF import LogPDF
from flowmc_amd.resource.model.nf_model.rqSpline import MaskedCouplingRQSpline
from flowmc_amd.resource.optimizer import AdamState, Optimizer, adamw
from flowmc_amd.resource.states import State
from flowmc_amd.strategy.take_steps import TakeGroupSteps
from flowmc_amd.strategy.train_model import TrainModel
from oracle import flow as F
from oracle import optim as O

pytestmark = pytest.mark.gpu

TOL = {"float32": 2e-4, "bfloat16": 2e-3}
CASES = [  # d, L, K, conditioner hidden widths
    (5, 4, 8, (32, 32)),
    (32, 3, 10, (128, 128)),
    (64, 2, 8, (256, 256)),
    (7, 3, 4, (16, 48)),
    (6, 2, 12, (32, 32)),  # K > 10: chunked spline path
    (6, 2, 8, (320, 64)),  # hidden > 256: the wide-hidden adjoint kernel
    (6, 3, 8, (48,)),  # one hidden layer
    (9, 2, 10, (64, 32, 48)),  # three hidden layers of different widths
    (16, 2, 8, (128, 128, 128, 128)),  # four hidden layers
    (8, 2, 12, (32, 32, 32)),  # three hidden layers, chunked spline path
    (6, 2, 7, (320,)),  # one wide hidden layer, runtime bin count
    (32, 8, 10, (128, 128)),  # the 32-D training flow at full depth
    (64, 10, 8, (256, 256)),  # the 64-D NF-MCMC flow at full depth
]
ADJ_CASES = CASES[:3] + CASES[4:]


def full_depth(case):
    return case[1] >= 8


def moderate(case):
    """Cases that draw moderate per-layer parameters: full depth and the
    conditioner-depth cases (the two-hidden-layer shallow cases keep the
    wilder parameters their bounds were set for)."""
    return full_depth(case) or len(case[3]) != 2


def bf16_q99(case):
    return 1e-2 if full_depth(case) else TOL["bfloat16"]


def bf16_max(case):
    return 3e-2 if moderate(case) else 0.15


def make(d, L, K, hidden, prec, seed=0, scale=None, hidden_var=1.0, out_var=None):
    hidden = list(hidden)
    mod = moderate((d, L, K, hidden))
    scale = (0.05 if mod else 0.3) if scale is None else scale
    out_var = (0.1 if mod else 1.0) if out_var is None else out_var
    p = F.init_params(d, L, K, hidden, seed=seed, scale=scale, hidden_var=hidden_var, out_var=out_var)
    g = torch.Generator().manual_seed(seed + 100)
    p["data_mean"] = torch.randn(d, generator=g, dtype=torch.float64) * 0.5
    A = torch.rand(d, generator=g, dtype=torch.float64) + 0.5
    p["data_cov"] = torch.diag(A) + 0.01 * torch.ones(d, d, dtype=torch.float64)  # off-diagonals are ignored
    p["base_cov"] = torch.eye(d, dtype=torch.float64) * 0.9
    p["base_mean"] = torch.full((d,), 0.1, dtype=torch.float64)
    theta = F.to_flat(p).float()
    p = F.from_flat(theta.double(), d, L, K, hidden)  # oracle on the FP32-rounded parameters
    m = MaskedCouplingRQSpline(d, L, hidden, K, fr.PRNGKey(0), precision=prec, _theta=theta)
    return m, p


def within(a, b, tol, what="", quantile=1.0):
    """scaled error |a-b|/(1+|b|) <= tol for the given quantile of elements."""
    a = a.detach().cpu().double()
    err = (a - b).abs() / (1 + b.abs())
    ok = bool(torch.quantile(err.reshape(-1), quantile) <= tol) if quantile < 1.0 else bool((err <= tol).all())
    if not ok:
        print(f"{what}: max scaled err {float(err.max()):.3e} (tol {tol:.1e}), "
              f"median {float(err.median()):.3e}, n>tol {int((err > tol).sum())}/{err.numel()}")
    return ok


@pytest.mark.parametrize("prec", ["float32", "bfloat16"])
@pytest.mark.parametrize("case", CASES)
def test_flow_passes_match_oracle(gpu, prec, case):
    """FP32: every element within 2e-4 of the FP64 oracle. BF16: against the
    oracle that rounds the GEMM operands to BF16 as the kernels do, the 99 %
    quantile and every element within the bounds of the module docstring."""
    d, L, K, H = case
    m, p = make(d, L, K, H, prec)
    g = torch.Generator().manual_seed(1)
    x = torch.randn(300, d, generator=g, dtype=torch.float64) * 3  # some samples land in the linear tails
    xf = x.float()
    tol = TOL[prec] if prec == "float32" else bf16_q99(case)
    q = 1.0 if prec == "float32" else 0.99
    y, ld = m.forward(xf.cuda())
    xi, ldi = m.inverse(xf.cuda())
    m.exact_inverse = True
    xe, _ = m.inverse(y.detach())
    m.exact_inverse = False
    z = torch.randn(200, d, generator=g)
    lp = m.log_prob(xf.cuda())
    smp = m.sample_from_noise(z.cuda())
    with (F.emulate_bf16() if prec == "bfloat16" else _null()):
        ry, rld = F.forward(p, xf.double(), K)
        rxi, rldi = F.inverse(p, xf.double(), K)
        rlp = F.log_prob(p, xf.double(), K)
        rsmp = F.sample_from_noise(p, z.double(), K)
    pairs = [("log_prob", lp, rlp, tol), ("forward y", y, ry, tol), ("forward logdet", ld, rld, tol),
             ("inverse x", xi, rxi, tol), ("inverse logdet", ldi, rldi, tol),
             ("exact inverse round trip", xe, xf.double(), 25 * tol), ("sample", smp, rsmp, tol)]
    checks = [within(a, b, t, what, q) for what, a, b, t in pairs]
    if prec == "bfloat16":
        checks += [within(a, b, bf16_max(case), what + " (all elements)") for what, a, b, _ in pairs]
    assert all(checks), checks


@pytest.mark.parametrize("prec", ["float32", "bfloat16"])
@pytest.mark.parametrize("case", CASES[:3] + CASES[6:9] + CASES[-2:])
def test_exact_inverse_samples_carry_their_log_prob(gpu, prec, case):
    """One exact-inverse pass returns samples and log N(z) - inverse log-det;
    that equals the forward log_prob of the same samples (change of
    variables) and the FP64 oracle's log_prob of them."""
    d, L, K, H = case
    m, p = make(d, L, K, H, prec, seed=6)
    m.exact_inverse = True
    g = torch.Generator().manual_seed(7)
    z = torch.randn(400, d, generator=g)
    x, lq = m.sample_and_log_prob(noise=z.cuda())
    xs = m.sample_from_noise(z.cuda())
    assert float((x - xs).abs().max()) == 0.0
    lf = m.log_prob(x)
    tol = TOL[prec] if prec == "float32" else bf16_q99(case)
    q = 1.0 if prec == "float32" else 0.99
    assert within(lq, lf.double().cpu(), 5 * tol, "inverse-pass log q vs forward log_prob", q)
    with (F.emulate_bf16() if prec == "bfloat16" else _null()):
        ref = F.log_prob(p, x.double().cpu(), K)
    assert within(lq, ref, 5 * tol, "inverse-pass log q vs oracle log_prob", q)
    m.exact_inverse = False
    with pytest.raises(ValueError):
        m.sample_and_log_prob(noise=z.cuda())


class _null:
    def __enter__(self):
        return None

    def __exit__(self, *a):
        return False


@pytest.mark.parametrize("case", CASES[:3] + CASES[6:9] + CASES[-2:])
def test_flow_bf16_error_vs_fp64_reference_init(gpu, case):
    """BF16 conditioner against the exact FP64 flow for parameters drawn like
    the reference's initialisation (hidden weights N(0, 1e-2/fan_in)): every
    log_prob within 1e-2 * (1 + |ref|)."""
    d, L, K, H = case
    m, p = make(d, L, K, H, "bfloat16", scale=0.05, hidden_var=1e-2)
    g = torch.Generator().manual_seed(11)
    x = (torch.randn(500, d, generator=g) * 2).float()
    assert within(m.log_prob(x.cuda()), F.log_prob(p, x.double(), K), 1e-2, "bf16 log_prob vs fp64")


@pytest.mark.parametrize("prec", ["float32", "bfloat16"])
@pytest.mark.parametrize("case", ADJ_CASES)
def test_flow_adjoint_matches_oracle_autograd(gpu, prec, case):
    d, L, K, H = case
    m, p = make(d, L, K, H, prec, seed=2)
    g = torch.Generator().manual_seed(3)
    B = 200
    x = (torch.randn(B, d, generator=g) * 2).float()
    w = torch.randn(B, generator=g).float()
    gt, gx = m.grad_log_prob_sum(x.cuda(), w.cuda())
    theta = F.to_flat(p).clone().requires_grad_(True)
    xr = x.double().clone().requires_grad_(True)
    with (F.emulate_bf16() if prec == "bfloat16" else _null()):
        obj = (w.double() * F.log_prob(F.from_flat(theta, d, L, K, list(H)), xr, K)).sum()
        rg, rgx = torch.autograd.grad(obj, [theta, xr])
    n_layers_part = m.n_layers * m._layer_size  # stat leaves get no gradient
    gt = gt.cpu().double()
    assert float(gt[n_layers_part:].abs().max()) == 0.0
    a, b = gt[:n_layers_part], rg[:n_layers_part]
    if prec == "float32":
        assert float((a - b).abs().max()) <= 2e-3 * float(b.abs().max())
        assert within(gx, rgx, 2e-3)
    else:
        # against the BF16-emulating oracle: relative L2 error of the whole gradient
        rel = float((a - b).norm() / b.norm())
        relx = float((gx.cpu().double() - rgx).norm() / rgx.norm())
        print(f"bf16 adjoint: rel L2 err theta {rel:.3e}, x {relx:.3e}")
        assert rel <= 1e-2 and relx <= 1e-2, (rel, relx)


@pytest.mark.parametrize("prec", ["float32", "bfloat16"])
@pytest.mark.parametrize("case,B", [(c, 200) for c in ADJ_CASES] + [((64, 2, 8, (256, 256)), 10000),
                                                                   ((32, 2, 10, (128, 128)), 10000)])
def test_flow_split_weight_gradient_matches_tile_kernel(gpu, prec, case, B):
    """The split-K weight gradient (batch ranges per workgroup, fixed-order
    sum of FP32 partials) against the one-workgroup-per-tile kernel: the same
    products of the same store fragments summed in another order, so equal to
    FP32 rounding (1e-5 of the largest gradient); the split one is also
    bit-identical run to run."""
    d, L, K, H = case
    m, _ = make(d, L, K, H, prec, seed=8)
    g = torch.Generator().manual_seed(9)
    x = (torch.randn(B, d, generator=g) * 2).float().cuda()
    w = (torch.randn(B, generator=g) / B).float().cuda()
    try:
        type(m).wgrad_split = False
        gt, gxt = m.grad_log_prob_sum(x, w)
        type(m).wgrad_split = True
        gs, gxs = m.grad_log_prob_sum(x, w)
        gs2, _ = m.grad_log_prob_sum(x, w)
    finally:
        type(m).wgrad_split = None
    assert torch.equal(gs, gs2)
    assert torch.equal(gxt, gxs)  # dL/dx does not involve the weight gradient
    n = m.n_layers * m._layer_size
    scale = float(gt[:n].abs().max())
    err = float((gs[:n] - gt[:n]).abs().max())
    assert err <= 1e-5 * scale, (err, scale)
    assert float(gs[n:].abs().max()) == 0.0


@pytest.mark.parametrize("prec", ["float32", "bfloat16"])
def test_flow_gradient_is_additive_over_batch_tiles(gpu, prec):
    """The gradient of a 40k-sample batch (1250 adjoint workgroups, several
    per CU, split-K weight gradients over 8 wavefronts) equals the sum of the
    gradients of its two halves, and dL/dx is per sample."""
    d, L, K, H = 64, 2, 8, (256, 256)
    m, _ = make(d, L, K, H, prec, seed=4)
    g = torch.Generator().manual_seed(5)
    B = 40000
    x = (torch.randn(B, d, generator=g) * 2).float().cuda()
    w = (torch.randn(B, generator=g) / B).float().cuda()
    gt, gx = m.grad_log_prob_sum(x, w)
    h = B // 2
    g1, gx1 = m.grad_log_prob_sum(x[:h].contiguous(), w[:h].contiguous())
    g2, gx2 = m.grad_log_prob_sum(x[h:].contiguous(), w[h:].contiguous())
    gs = (g1 + g2).double()
    n = m.n_layers * m._layer_size
    err = float((gt.double()[:n] - gs[:n]).abs().max() / gs[:n].abs().max())
    assert err <= 1e-4, err
    gxs = torch.cat([gx1, gx2])
    assert float((gx - gxs).abs().max()) <= 1e-6 * float(gxs.abs().max())


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("hidden", [128, 256])
def test_flow_gradient_independent_of_the_backward_schedule(gpu, hidden, split):
    """Weight gradients on a side stream beside the next adjoint (the default)
    or on one stream: the same kernels on the same data, so the same bits
    (each weight-gradient kernel in turn)."""
    d, L, K = 32, 4, 10
    m, _ = make(d, L, K, (hidden, hidden), "bfloat16", seed=6)
    type(m).wgrad_split = split
    try:
        _schedule_bits(m, d, L, hidden)
    finally:
        type(m).wgrad_split = None


@pytest.mark.parametrize("split", [False, True])
def test_flow_backward_clears_the_gradient(gpu, split):
    """fmc_flow_backward clears grad before the weight gradients: a
    workspace gradient left holding NaN gives the same bits as a fresh one
    (masked input columns and the data statistics are never written)."""
    d, L, K, B = 8, 3, 6, 700
    m, _ = make(d, L, K, (32, 32), "bfloat16", seed=9)
    type(m).wgrad_split = split
    try:
        g = torch.Generator().manual_seed(10)
        x = (torch.randn(B, d, generator=g) * 2).float().cuda()
        w = (torch.randn(B, generator=g) / B).float().cuda()
        g1, gx1 = m.grad_log_prob_sum(x, w)
        m._ws[B].grad.fill_(float("nan"))
        g2, gx2 = m.grad_log_prob_sum(x, w)
        assert not torch.isnan(g2).any()
        assert torch.equal(g1, g2) and torch.equal(gx1, gx2)
    finally:
        type(m).wgrad_split = None


def _schedule_bits(m, d, L, hidden):
    g = torch.Generator().manual_seed(7)
    B = 10000
    x = (torch.randn(B, d, generator=g) * 2).float().cuda()
    w = (torch.randn(B, generator=g) / B).float().cuda()
    g1, gx1 = m.grad_log_prob_sum(x, w)
    ws = m._ws[B]
    assert ws.overlap  # the default schedule; then the same launches on one stream
    ws.overlap = False
    for _ in range(2):
        g2, gx2 = m.grad_log_prob_sum(x, w)
        assert torch.equal(g1, g2) and torch.equal(gx1, gx2)


@pytest.mark.parametrize("case", CASES[:3] + CASES[6:9] + CASES[-2:])
def test_flow_bf16_gradient_vs_fp64_reference_init(gpu, case):
    """BF16 adjoint against the exact FP64 gradient at reference-like init."""
    d, L, K, H = case
    m, p = make(d, L, K, H, "bfloat16", seed=2, scale=0.05, hidden_var=1e-2)
    g = torch.Generator().manual_seed(3)
    x = (torch.randn(400, d, generator=g) * 2).float()
    gt, _ = m.grad_log_prob_sum(x.cuda(), torch.full((400,), -1.0 / 400).cuda())
    theta = F.to_flat(p).clone().requires_grad_(True)
    (rg,) = torch.autograd.grad(F.loss(F.from_flat(theta, d, L, K, list(H)), x.double(), K), theta)
    n = m.n_layers * m._layer_size
    a, b = gt.cpu().double()[:n], rg[:n]
    cos = float((a * b).sum() / (a.norm() * b.norm()))
    rel = float((a - b).norm() / b.norm())
    print(f"bf16 gradient vs fp64: cos {cos:.6f}, rel L2 {rel:.3e}")
    assert cos >= 0.999 and rel <= 5e-2, (cos, rel)


@pytest.mark.parametrize("n", [100003, 3])
@pytest.mark.parametrize("clip", [True, False])
def test_adamw_kernel_matches_oracle(gpu, clip, n):
    """Clip + AdamW against the FP64 oracle; n = 3 runs only the kernels'
    scalar tails (fewer than one float4)."""
    g = torch.Generator().manual_seed(4)
    theta = torch.randn(n, generator=g)
    th, mm, vv = theta.double(), torch.zeros(n, dtype=torch.float64), torch.zeros(n, dtype=torch.float64)
    dt, dm, dv = theta.cuda(), torch.zeros(n, device=gpu), torch.zeros(n, device=gpu)
    ws = torch.empty(1024, dtype=torch.float64, device=gpu)
    gn = torch.empty(1, device=gpu)
    for t in range(1, 4):
        grad = torch.randn(n, generator=g) * (0.05 if clip else 1e-4)
        th, mm, vv = O.adamw_clip_step(th, grad.double(), mm, vv, t, 1e-3, max_norm=1.0)
        gd = grad.cuda()
        N.call("fmc_adamw_clip", N.stream_ptr(), dt.data_ptr(), gd.data_ptr(), dm.data_ptr(), dv.data_ptr(), n, 1e-3,
               0.9, 0.999, 1e-8, 1e-4, 1.0, t, ws.data_ptr(), 1024, gn.data_ptr())
        assert abs(float(gn) - float(grad.double().norm())) <= 1e-5 * float(grad.double().norm())
    assert float((dt.cpu().double() - th).abs().max()) <= 1e-6


@pytest.mark.parametrize("hidden", [(32, 32), (128, 128)])
def test_graph_train_steps_match_eager_steps(gpu, hidden):
    """Training steps replayed from a recorded HIP graph (graph_train) give
    the eager steps' loss, parameters and Adam moments bit for bit, step
    after step, with new batches and eager work between the replays; the
    model then evaluates with the updated parameters."""
    d, L, K = 8, 3, 8
    m1 = MaskedCouplingRQSpline(d, L, list(hidden), K, fr.PRNGKey(0))
    m2 = m1.clone()  # (the clone shares m1's workspace: their steps interleave on it)
    m1.graph_train = False
    m2.graph_train = True
    opt = adamw(1e-3, max_norm=1.0)
    s1, s2 = opt.init(m1), opt.init(m2)
    g = torch.Generator().manual_seed(8)
    for step in range(4):
        xb = (torch.randn(512, d, generator=g) * 2).cuda()
        l1, m1, s1 = m1.train_step(xb, opt, s1)
        l2, m2, s2 = m2.train_step(xb, opt, s2)
        _ = float(m2.log_prob(xb[:16]).sum())  # eager work (and a repack) between replays
        torch.cuda.synchronize()
        assert torch.equal(l1, l2), step
        assert torch.equal(m1.theta, m2.theta), step
        assert torch.equal(s1.mu, s2.mu) and torch.equal(s1.nu, s2.nu), step
        assert s1.count == s2.count == step + 1
    x = torch.randn(64, d, generator=g).cuda()
    assert torch.equal(m1.log_prob(x), m2.log_prob(x))


def test_train_step_matches_oracle_fp32(gpu):
    d, L, K, H = 6, 2, 5, (32, 32)
    m, p = make(d, L, K, H, "float32", seed=5)
    g = torch.Generator().manual_seed(6)
    x = (torch.randn(256, d, generator=g) * 1.5).float()
    opt = adamw(1e-3, weight_decay=1e-4, max_norm=1.0)
    state = opt.init(m)
    loss, m2, state = m.train_step(x.cuda(), opt, state)
    theta = F.to_flat(p).clone().requires_grad_(True)
    rl = F.loss(F.from_flat(theta, d, L, K, list(H)), x.double(), K)
    (rg,) = torch.autograd.grad(rl, theta)
    th, _, _ = O.adamw_clip_step(theta.detach(), rg, torch.zeros_like(rg), torch.zeros_like(rg), 1, 1e-3,
                                 max_norm=1.0)
    assert abs(float(loss) - float(rl.detach())) <= 2e-4 * (1 + abs(float(rl.detach())))
    # the first Adam step moves every parameter by ~lr*sign(g): compare the update where |g| is not tiny
    upd = m2.theta.cpu().double() - F.to_flat(p)
    rupd = th - F.to_flat(p)
    big = rg.abs() > 1e-3 * rg.abs().max()
    assert float((upd[big] - rupd[big]).abs().max()) <= 2e-5
    assert state.count == 1


def test_train_keeps_the_first_lowest_loss_epoch(gpu):
    """NFModel.train picks the epoch of lowest last-minibatch loss on the
    device (base.py:196-200: the first such epoch); the same loop with a host
    read-back of every epoch's loss and a clone at each improvement, as the
    reference writes it, returns the same parameters, Adam moments and count
    and the same loss values."""
    from flowmc_amd.resource.model.nf_model.base import NFModel, data_stats

    d = 4
    g = torch.Generator().manual_seed(17)
    data = (torch.randn(3000, d, generator=g) * 1.5).cuda()
    m = MaskedCouplingRQSpline(d, 3, [32, 32], 8, fr.PRNGKey(4))
    opt = Optimizer(m, learning_rate=3e-2)  # large steps: the loss goes up and down
    rng, best, st, losses = m.train(fr.PRNGKey(5), data, opt.optim, opt.optim_state, 6, 500, verbose=False)
    # the reference's loop
    model, state = m.clone(), opt.optim_state.clone()
    mean, cov = data_stats(data)  # train's jnp.mean / jnp.cov (ddof 1), as torch gives them
    assert torch.allclose(cov, torch.cov(data.T), rtol=1e-5, atol=1e-6)
    assert torch.allclose(mean, data.mean(0), rtol=0, atol=0)
    model.set_data_stats(mean, cov)
    key, ref_best, ref_loss, ref_losses = fr.PRNGKey(5), None, 1e9, []
    for _ in range(6):
        key, sub = fr.split(key)
        value, model, state = NFModel.train_epoch(model, sub, opt.optim, state, data, 500)
        v = float(value)
        ref_losses.append(v)
        if v < ref_loss:
            ref_best, ref_loss = (model.clone(), state.clone()), v
    assert torch.equal(losses, torch.tensor(ref_losses))
    assert torch.equal(best.theta, ref_best[0].theta)
    assert st.count == ref_best[1].count
    assert torch.equal(st.mu, ref_best[1].mu) and torch.equal(st.nu, ref_best[1].nu)


@pytest.mark.parametrize("graph", [False, True])
def test_train_epoch_matches_a_loop_of_train_steps(gpu, graph):
    """train_epoch skips the loss reduction on all but its last minibatch
    (base.py:127-151 returns only that loss); parameters, Adam moments, step
    count and the returned loss equal a plain loop of train_step over the
    same permutation, bit for bit."""
    from flowmc_amd.resource.model.nf_model.base import NFModel

    d, bs = 4, 500
    g = torch.Generator().manual_seed(23)
    data = (torch.randn(2600, d, generator=g) * 1.3).cuda()
    m = MaskedCouplingRQSpline(d, 3, [32, 32], 8, fr.PRNGKey(8))
    m.graph_train = graph
    opt = Optimizer(m, learning_rate=1e-2)
    a, sa = m.clone(), opt.optim_state.clone()
    value, a, sa = NFModel.train_epoch(a, fr.PRNGKey(9), opt.optim, sa, data, bs)
    b, sb = m.clone(), opt.optim_state.clone()
    steps = data.shape[0] // bs
    perms = fr.permutation(fr.PRNGKey(9), data.shape[0], device=data.device)[: steps * bs].reshape(steps, bs)
    for perm in perms:
        vb, b, sb = b.train_step(data[perm], opt.optim, sb)
    assert sa.count == sb.count == steps
    assert torch.equal(value, vb)
    assert torch.equal(a.theta, b.theta)
    assert torch.equal(sa.mu, sb.mu) and torch.equal(sa.nu, sb.nu)


def test_train_rounds_recorded_match_eager(gpu):
    """Three train() calls in a row, each on the previous call's best model
    (a sampling run's training rounds): with recorded steps (one recording
    reused across the calls through the shared working copy) and eagerly,
    the same losses, parameters and Adam moments bit for bit; the caller's
    models are left unchanged."""
    d = 4
    g = torch.Generator().manual_seed(31)
    data = (torch.randn(2600, d, generator=g) * 1.2).cuda()
    out = {}
    for graph in (False, True):
        m = MaskedCouplingRQSpline(d, 3, [32, 32], 8, fr.PRNGKey(12))
        m.graph_train = graph
        opt = Optimizer(m, learning_rate=1e-2)
        model, state, key, hist = m, opt.optim_state, fr.PRNGKey(13), []
        for _ in range(3):
            before = model.theta.clone()
            key, model2, state, losses = model.train(key, data, opt.optim, state, 3, 500, verbose=False)
            assert torch.equal(model.theta, before)
            assert model2 is not model or bool((losses >= 1e9).all())
            model = model2
            hist.append(losses)
        if graph:
            assert len(m._slots) == 1 and getattr(m._ws[500], "graph", None) is not None
        out[graph] = (torch.stack(hist), model.theta.clone(), state.mu.clone(), state.nu.clone(), state.count)
    a, b = out[False], out[True]
    assert torch.equal(a[0], b[0])
    assert torch.equal(a[1], b[1]) and torch.equal(a[2], b[2]) and torch.equal(a[3], b[3]) and a[4] == b[4]


def test_training_reduces_loss_and_best_model(gpu):
    d = 4
    g = torch.Generator().manual_seed(7)
    means = torch.tensor([[2.0, -1.0, 0.0, 1.0], [-2.0, 1.0, 1.0, -1.0]])
    lab = torch.randint(0, 2, (4000,), generator=g)
    data = (means[lab] + 0.5 * torch.randn(4000, d, generator=g)).cuda()
    m = MaskedCouplingRQSpline(d, 4, [32, 32], 8, fr.PRNGKey(1))
    opt = Optimizer(m, learning_rate=1e-2)
    rng, best, state, losses = m.train(fr.PRNGKey(2), data, opt.optim, opt.optim_state, 8, 500, verbose=False)
    assert losses.shape == (8,)
    assert float(losses[-1]) < float(losses[0])
    assert float(losses.min()) == pytest.approx(float(best.loss_fn(data)), abs=0.5)
    # the caller's model is unchanged (functional train)
    assert float(m.data_mean.abs().max()) == 0.0 and float(best.data_mean.abs().max()) > 0.0


def test_nf_proposal_accepts_all_when_target_is_the_flow(gpu):
    # test/unit/test_strategies.py:251-319 with the assertion the reference leaves out
    n_chains, n_steps, d = 5, 25, 2
    model = MaskedCouplingRQSpline(d, 3, [16, 16], 8, fr.PRNGKey(10))
    resources = {
        "test_position": Buffer("test_position", (n_chains, n_steps, d), 1),
        "test_log_prob": Buffer("test_log_prob", (n_chains, n_steps), 1),
        "test_acceptance": Buffer("test_acceptance", (n_chains, n_steps), 1),
        "NFProposal": NFProposal(model, n_NFproposal_batch_size=5),
        "logpdf": LogPDF(lambda x, data: model.log_prob(x), n_dims=d),
        "sampler_state": State({"test_position": "test_position", "test_log_prob": "test_log_prob",
                                "test_acceptance": "test_acceptance"}, name="sampler_state"),
    }
    x0 = torch.zeros(n_chains, d, device=gpu)
    for S in (11, 5):
        strat = TakeGroupSteps("logpdf", "NFProposal", "sampler_state",
                               ["test_position", "test_log_prob", "test_acceptance"], n_steps=S)
        strat(fr.PRNGKey(42), resources, x0, {"data": torch.arange(d)})
        acc = resources["test_acceptance"].data[:, :S]
        assert bool((acc == 1).all())
        pos = resources["test_position"].data[:, :S]
        assert torch.allclose(resources["test_log_prob"].data[:, :S], model.log_prob(pos), atol=1e-5)


def test_train_model_strategy(gpu):
    # test/unit/test_strategies.py:224-249 plus checks of the data selection
    n_chains, n_steps, d = 5, 25, 2
    model = MaskedCouplingRQSpline(d, 3, [16, 16], 8, fr.PRNGKey(10))
    buf = Buffer("test_data", (n_chains, n_steps, d), 1)
    buf.update_buffer(fr.normal(fr.PRNGKey(0), (n_chains, n_steps - 5, d)))  # last 5 rows stay -inf
    opt = Optimizer(model)
    res = {"test_data": buf, "optimizer": opt, "model": model,
           "loss": Buffer("loss", (10,), 0)}
    strat = TrainModel("model", "test_data", "optimizer", loss_buffer_name="loss", n_epochs=10,
                       batch_size=n_chains * n_steps, n_max_examples=10000, history_window=1)
    _, res, _ = strat(fr.PRNGKey(42), res, torch.zeros(n_chains, d, device=gpu), {})
    new = res["model"]
    assert isinstance(new, MaskedCouplingRQSpline) and new is not model
    # history_window=1: training data are the last finite row of each chain
    last = buf.data[:, n_steps - 6]
    assert torch.allclose(new.data_mean, last.mean(0), atol=5e-2)
    assert bool(torch.isfinite(res["loss"].data).all()) and res["loss"].cursor == 10
    assert opt.optim_state.count >= 1


def test_flow_log_prob_autograd_function(gpu):
    d, L, K, H = 5, 3, 6, (32, 32)
    m, p = make(d, L, K, H, "float32", seed=8)
    x = (torch.randn(40, d) * 1.5).cuda().requires_grad_(True)
    lp = m.log_prob(x)
    (gx,) = torch.autograd.grad(lp.sum(), x)
    xr = x.detach().cpu().double().requires_grad_(True)
    (rgx,) = torch.autograd.grad(F.log_prob(p, xr, K).sum(), xr)
    assert within(gx, rgx, 2e-3, "d log_prob / dx")


def test_mala_on_a_flow_target(gpu):
    # a user target that calls the flow: LogPDF falls back to the batched path
    from flowmc_amd.resource.kernel.base import StepOutputs
    from flowmc_amd.resource.kernel.MALA import MALA

    d = 4
    m = MaskedCouplingRQSpline(d, 2, [32, 32], 8, fr.PRNGKey(3), precision="float32")
    lpdf = LogPDF(lambda x, data: m.log_prob(x), n_dims=d)
    C_, S = 64, 20
    pos = torch.zeros(C_, S, d, device=gpu)
    lp = torch.zeros(C_, S, device=gpu)
    ac = torch.zeros(C_, S, device=gpu)
    MALA(0.5).sample_steps(fr.PRNGKey(1), torch.zeros(C_, d, device=gpu), lpdf, None, S, 1,
                           StepOutputs(pos, lp, ac, 0))
    assert bool(torch.isfinite(pos).all()) and 0.2 < float(ac.mean()) <= 1.0
    assert torch.allclose(lp, m.log_prob(pos), atol=1e-4)


@pytest.mark.parametrize("hidden", [[24], [32, 16, 32], [16, 16, 16, 16]])
def test_training_with_other_conditioner_depths_reduces_loss(gpu, hidden):
    """hidden_size lists of one, three and four layers train end to end
    (rqSpline.py:428-433 builds the MLP from any list)."""
    d = 4
    g = torch.Generator().manual_seed(17)
    means = torch.tensor([[2.0, -1.0, 0.0, 1.0], [-2.0, 1.0, 1.0, -1.0]])
    lab = torch.randint(0, 2, (4000,), generator=g)
    data = (means[lab] + 0.5 * torch.randn(4000, d, generator=g)).cuda()
    m = MaskedCouplingRQSpline(d, 3, hidden, 8, fr.PRNGKey(4))
    mlp = m.layers[0][1].bijector
    assert [tuple(lin.weight.shape) for lin in mlp.layers if hasattr(lin, "weight")] == [
        (o, i) for i, o in zip([d] + hidden, hidden + [d * (3 * 8 + 1)])]
    opt = Optimizer(m, learning_rate=1e-2)
    rng, best, state, losses = m.train(fr.PRNGKey(2), data, opt.optim, opt.optim_state, 8, 500, verbose=False)
    assert float(losses[-1]) < float(losses[0])


def test_deep_wide_conditioner_adjoint_is_refused(gpu):
    """Three hidden layers wider than 256: the pass runs, the adjoint kernel
    refuses them with a message (not a silent fallback)."""
    m = MaskedCouplingRQSpline(6, 2, [320, 64, 64], 8, fr.PRNGKey(3))
    x = torch.randn(64, 6, device=gpu)
    assert bool(torch.isfinite(m.log_prob(x)).all())
    with pytest.raises(N.FlowMCNativeError, match="at most two hidden layers"):
        m.grad_log_prob_sum(x)


@pytest.mark.parametrize("prec", ["float32", "bfloat16"])
@pytest.mark.parametrize("case", [(6, 3, 8, (32, 32)), (16, 2, 10, (64, 32, 64))])
def test_dense_base_covariance_matches_oracle(gpu, prec, case):
    """A base Gaussian with a full covariance (Gaussian.log_prob / sample,
    common.py:285-293): log_prob, samples (mean + chol(cov) xi), the exact
    inverse's own density and the adjoint against the FP64 oracle, which
    evaluates multivariate_normal with the full covariance."""
    d, L, K, H = case
    p = F.init_params(d, L, K, list(H), seed=9, scale=0.05, out_var=0.1)
    g = torch.Generator().manual_seed(10)
    A = torch.randn(d, d, generator=g, dtype=torch.float64)
    p["base_cov"] = A @ A.T / d + 0.5 * torch.eye(d, dtype=torch.float64)
    p["base_mean"] = 0.3 * torch.randn(d, generator=g, dtype=torch.float64)
    p["data_mean"] = 0.2 * torch.randn(d, generator=g, dtype=torch.float64)
    p["data_cov"] = torch.diag(0.5 + torch.rand(d, generator=g, dtype=torch.float64))
    theta = F.to_flat(p).float()
    p = F.from_flat(theta.double(), d, L, K, list(H))
    m = MaskedCouplingRQSpline(d, L, list(H), K, fr.PRNGKey(0), precision=prec, _theta=theta)
    assert m._base_dense
    tol = TOL[prec]
    q = 1.0 if prec == "float32" else 0.99
    x = torch.randn(256, d, generator=g).float() * 1.5
    z = torch.randn(256, d, generator=g).float()
    lp = m.log_prob(x.cuda())
    smp = m.sample_from_noise(z.cuda())
    m.exact_inverse = True
    xs, lq = m.sample_and_log_prob(noise=z.cuda())
    m.exact_inverse = False
    with (F.emulate_bf16() if prec == "bfloat16" else _null()):
        rlp = F.log_prob(p, x.double(), K)
        rsmp = F.sample_from_noise(p, z.double(), K)
        rxs = F.sample_from_noise(p, z.double(), K, exact=True)
        rlq = F.log_prob(p, rxs, K)
    checks = [within(lp, rlp, tol, "log_prob", q), within(smp, rsmp, tol, "sample", q),
              within(xs, rxs, tol, "exact sample", q), within(lq, rlq, 5 * tol, "exact sample log q", q)]
    assert all(checks), checks
    # adjoint: d/dtheta and d/dx of sum w log_prob (the base statistics are stop-gradient leaves)
    w = torch.randn(256, generator=g).float()
    gt, gx = m.grad_log_prob_sum(x.cuda(), w.cuda())
    th = F.to_flat(p).clone().requires_grad_(True)
    xr = x.double().clone().requires_grad_(True)
    with (F.emulate_bf16() if prec == "bfloat16" else _null()):
        obj = (w.double() * F.log_prob(F.from_flat(th, d, L, K, list(H)), xr, K)).sum()
        rg, rgx = torch.autograd.grad(obj, [th, xr])
    n = m.n_layers * m._layer_size
    a, b = gt.cpu().double()[:n], rg[:n]
    if prec == "float32":
        assert float((a - b).abs().max()) <= 2e-3 * float(b.abs().max())
        assert within(gx, rgx, 2e-3, "dx")
    else:
        assert float((a - b).norm() / b.norm()) <= 1e-2
        assert float((gx.cpu().double() - rgx).norm() / rgx.norm()) <= 1e-2


def test_learnable_base_gradient_matches_oracle(gpu):
    """Gaussian(learnable=True) base (common.py:243-273): the flat gradient of
    sum w log_prob carries the base mean and covariance entries, against the
    FP64 oracle's autograd with the learnable base (the layer entries as with
    a frozen base, the data statistics zero); a training step moves the base
    and keeps its covariance symmetric and positive definite."""
    from flowmc_amd.resource.model.common import Gaussian
    from flowmc_amd.resource.optimizer import Optimizer

    d, L, K, H = 6, 3, 8, [32, 32]
    p = F.init_params(d, L, K, H, seed=9, scale=0.05, out_var=0.1)
    g = torch.Generator().manual_seed(31)
    A = torch.randn(d, d, generator=g, dtype=torch.float64)
    p["base_cov"] = A @ A.T / d + 0.5 * torch.eye(d, dtype=torch.float64)
    p["base_mean"] = 0.3 * torch.randn(d, generator=g, dtype=torch.float64)
    p["data_mean"] = 0.2 * torch.randn(d, generator=g, dtype=torch.float64)
    p["data_cov"] = torch.diag(0.5 + torch.rand(d, generator=g, dtype=torch.float64))
    theta = F.to_flat(p).float()
    p = F.from_flat(theta.double(), d, L, K, H)
    bd = Gaussian(p["base_mean"].float(), p["base_cov"].float(), learnable=True)
    m = MaskedCouplingRQSpline(d, L, H, K, fr.PRNGKey(0), precision="float32", _theta=theta, base_dist=bd)
    assert m._base_dense and m.base_dist.learnable
    x = torch.randn(256, d, generator=g).float() * 1.5
    w = torch.randn(256, generator=g).float()
    gt, gx = m.grad_log_prob_sum(x.cuda(), w.cuda())
    th = F.to_flat(p).clone().requires_grad_(True)
    xr = x.double().clone().requires_grad_(True)
    obj = (w.double() * F.log_prob(F.from_flat(th, d, L, K, H), xr, K, learnable_base=True)).sum()
    rg, rgx = torch.autograd.grad(obj, [th, xr])
    gt = gt.cpu().double()
    n = m.n_layers * m._layer_size
    om, oc = m._off["base_mean"], m._off["base_cov"]
    assert float((gt[:n] - rg[:n]).abs().max()) <= 2e-3 * float(rg[:n].abs().max())
    for lo, hi in ((om, om + d), (oc, oc + d * d)):
        a, b = gt[lo:hi], rg[lo:hi]
        assert float((a - b).abs().max()) <= 2e-3 * float(b.abs().max()), (lo, a, b)
    assert float(gt[n:om].abs().max()) == 0.0  # data statistics stay frozen
    assert within(gx, rgx, 2e-3, "dx")
    # a training step (clip + AdamW over every entry) moves the learnable base
    opt = Optimizer(m, learning_rate=1e-2)
    mean0, cov0 = m.base_dist.mean.clone(), m.base_dist.cov.clone()
    loss, m2, state = m.train_step(x.cuda(), opt.optim, opt.optim_state)
    cov1 = m2.base_dist.cov
    assert float((m2.base_dist.mean - mean0).abs().max()) > 1e-4
    assert float((cov1 - cov0).abs().max()) > 1e-4
    assert torch.allclose(cov1, cov1.T, atol=1e-6)
    assert bool(torch.isfinite(torch.linalg.cholesky(cov1.double())).all())
    assert bool(torch.isfinite(m2.log_prob(x.cuda())).all())
