"""Benchmarks of the flowMC hot path on one MI355X. Prints ONE JSON line.

Workloads (``--workload``):
  nfmcmc_gmm64  (default) full NF-MCMC loop (RQSpline_MALA bundle through
                Sampler) on the 64-D 8-mode Gaussian mixture: 4096 chains,
                MALA step 0.05, RQ-spline flow 10 layers / 8 bins / [256, 256]
                with BF16 conditioner GEMMs, ``--steps`` training rounds then
                ``--steps`` production rounds of 100 local + 100 global steps;
                training per round: ``--epochs`` (3) epochs, batch 10000,
                lr 2e-4, 100k examples from the last 200 rows of every chain;
                chains start from one exact mixture draw each (a second run
                from x0 ~ N(0, I) is reported in ``extra.cold_start``).
                value = min bulk-ESS over dims of the production samples /
                end-to-end wall time (training included); valid when max
                R-hat < 1.01. The same workload run by the plain FP32
                per-step GPU path (bench_plain.py) is reported beside it.
  mala_rosenbrock_128d  fused MALA on the 128-D Rosenbrock target, 16384
                chains, step 0.01, 64 steps per launch; one step = one MALA
                step of all chains (chain-steps/s).
  flow_train_d32  flow training only: d=32, 8 layers, 10 bins, [128, 128],
                BF16 conditioner, 100k samples of a 4-component mixture,
                batch 10k, Adam lr 1e-3 (training samples/s).
  hmc_funnels   32 Neal's funnels of log-uniform dimension in [8, 256], 256
                chains each, 10 leapfrog steps of 0.05, one batched launch per
                16 steps (chain-steps/s).
  grw_rosenbrock_128d  fused Gaussian random walk, same target / chains /
                step (chain-steps/s).
  flow_pass     the exact-inverse flow pass over the 409,600 proposals of one
                global step of nfmcmc_gmm64 (samples/s; profiling target).
  flow_adjoint  one coupling layer's adjoint at the training shape of
                nfmcmc_gmm64 (d=64, [256, 256], batch 10k; samples/s).
  mala_gauss5d  fused MALA on the 5-D correlated Gaussian, 1000 chains, step
                0.1, 1000 steps from x0 ~ N(0, I), then one flow training
                round (4 layers, 8 bins, [32, 32], batch 1000, 5 epochs).
  realnvp_train_d32  RealNVP training (d=32, 8 layers, hidden 256, batch 10k)
                and a 409,600-sample draw.
  pt_mala_gmm64  parallel tempering (5 temperatures) of fused MALA on the
                64-D mixture, 4096 chains (tempered chain-steps/s).
``--gpus N`` with N > 1 exits: the build runs on a single GPU.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP32_PEAK_TFLOPS = 157.3


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def flow_layer_flops(d, layer, K, hidden):
    """Minimal conditioner FLOPs of one coupling layer per sample:
    2 * (d_c*h_0 + sum_i h_{i-1}*h_i + h_last*d_t*(3K+1))."""
    dc = (d + 1) // 2 if layer % 2 else d // 2
    dt = d - dc
    tot = dc * hidden[0] + sum(a * b for a, b in zip(hidden[:-1], hidden[1:])) + hidden[-1] * dt * (3 * K + 1)
    return 2 * tot


def flow_wgrad_macs(d, layer, K, hidden):
    """Multiply-adds per sample of one coupling layer's weight gradients
    (dW = G^T A for every linear layer: rows x cols each, the same count as
    the layer's forward product)."""
    return flow_layer_flops(d, layer, K, hidden) // 2


def flow_flops_per_sample(d, L, K, h1, h2=None):
    """Minimal conditioner FLOPs of one pass (hidden widths (h1, h2) or the
    list h1): the sum of flow_layer_flops over the layers."""
    hidden = [h1, h2] if h2 is not None else list(h1)
    return sum(flow_layer_flops(d, l, K, hidden) for l in range(L))


def cpu_rate(make_work, n_all, n_one=None, reps=5, warmup=1):
    """Throughput of a CPU leg. ``make_work(n)`` returns a callable that does n
    units of work (chain-steps, samples, ...). With all host threads on n_all
    units and with one thread on n_one units (n_all // 8 by default): one
    warm-up call, then the median time of ``reps`` calls. Returns
    (units/s all threads, units/s one thread, threads used)."""
    import statistics

    import torch

    threads = os.cpu_count() or 1
    n_one = n_one or max(1, n_all // 8)
    rates = []
    for nt, n in ((threads, n_all), (1, n_one)):
        torch.set_num_threads(nt)
        fn = make_work(n)
        for _ in range(warmup):
            fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        rates.append(n / statistics.median(ts))
    torch.set_num_threads(threads)
    return rates[0], rates[1], threads


def trace_marker():
    """One tiny fmc_philox_blocks launch: in a kernel trace, the launches
    between two of them are the timed region (nothing else launches that
    kernel in these workloads)."""
    import torch

    import flowmc_amd._native as N

    buf = torch.empty(4, dtype=torch.int32, device="cuda")
    N.call("fmc_philox_blocks", N.stream_ptr(), 0, 0, 0, 0, 0, 0, 1, buf.data_ptr())


def mode_occupancy(pos, target):
    """Fraction of the samples [..., d] whose most likely mixture component
    (largest weighted component density) is each mode."""
    import torch

    x = pos.reshape(-1, pos.shape[-1])
    mu = target.means.to(x.device, torch.float32)
    sg = target.sigmas.to(x.device, torch.float32)
    lw = torch.log(target.weights.to(x.device, torch.float32))
    counts = torch.zeros(mu.shape[0], dtype=torch.float64, device=x.device)
    for i in range(0, x.shape[0], 65536):
        xb = x[i:i + 65536]
        z = (xb[:, None, :] - mu) / sg
        q = lw - torch.log(sg).sum(-1) - 0.5 * (z * z).sum(-1)
        counts += torch.bincount(q.argmax(-1), minlength=mu.shape[0]).double()
    return (counts / counts.sum()).tolist()


def time_kernel(fn, reps=5):
    """Average time (ms) of fn() measured with HIP events on the current stream."""
    import torch

    fn()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


# --------------------------------------------------------------- C5 -------
def nfmcmc_config(args):
    """The 64-D 8-mode NF-MCMC workload. Fixed by the workload: 4096 chains,
    MALA step 0.05, RQ-spline 10 layers / 8 bins / [256, 256], 100 local +
    100 global steps per round, ``--steps`` training and production rounds.
    Chosen here (the same for every path): 3 epochs per round over 100k
    examples drawn from the last 200 rows of every chain (this round's MALA
    rows and the previous round's NF rows; MALA at 0.05 leaves the 100 local
    rows near-copies of one point, so a 100-row window trains on only ~4096
    distinct points per round), batch 10k, AdamW lr 2e-4 (more epochs over
    the same slowly-moving chain clouds overfit the 4.9M parameters and the
    global acceptance collapses), and sampling with the exact inverse of the
    forward map."""
    return dict(n_local=100, n_global=100, n_loops=args.steps, n_epochs=args.epochs, batch_size=args.batch,
                n_max_examples=args.nmax, step_size=0.05, lr=args.lr, n_layers=10, n_bins=8, hidden=[256, 256],
                exact_inverse=True, history_window=args.window)


def nfmcmc_initial_positions(target, n_chains, device, seed=1):
    """One exact draw of the mixture per chain (mode uniformly at random, then
    the mode's diagonal Gaussian): the state a perfect burn-in would reach, so
    the run measures how the sampler keeps and mixes the 8 modes. From N(0, I)
    MALA at step 0.05 needs thousands of steps to reach any mode and the
    chains start in one basin, which no global step can rebalance."""
    import torch

    g = torch.Generator(device=device).manual_seed(seed)
    mu = target.means.to(device, torch.float32)
    sg = target.sigmas.to(device, torch.float32)
    lab = torch.randint(0, mu.shape[0], (n_chains,), device=device, generator=g)
    return (mu[lab] + sg[lab] * torch.randn(n_chains, mu.shape[1], device=device, generator=g)).contiguous()


def nfmcmc_analysis(pos, target, n_local, n_global):
    """Diagnostics of the production buffer [C, rounds*(n_local+n_global), d].
    Analysis view: each local block is represented by its last row, every
    global row is kept. MALA at step 0.05 barely moves in 100 steps, so the
    100 near-copies of a local block add no information but dominate the
    within-chain autocorrelation: on the full buffer even an exact global
    sampler gives a split-R-hat of ~1.012 at 4096 chains. Validity: R-hat of
    the view < 1.01, every |mean error| <= 1e-2 sd and every relative variance
    error <= 1e-2."""
    import torch

    from flowmc_amd import diagnostics as D

    per = n_local + n_global
    rounds = pos.shape[1] // per
    idx = []
    for r in range(rounds):
        idx.append(r * per + n_local - 1)
        idx.extend(range(r * per + n_local, (r + 1) * per))
    view = pos[:, torch.tensor(idx, device=pos.device)]
    sv = D.summary(view)
    full_rhat = float(D.rhat(pos).max())
    dev = pos.device
    mean_t, var_t = target.mean().to(dev), target.variance().to(dev)
    mean_err = float(((sv["mean"] - mean_t).abs() / var_t.sqrt()).max())
    var_err = float(((sv["var"] - var_t) / var_t).abs().max())
    valid = sv["max_rhat"] < 1.01 and mean_err <= 1e-2 and var_err <= 1e-2
    return {"min_ess_bulk": sv["min_ess_bulk"], "max_rhat": sv["max_rhat"], "max_rhat_full_buffer": full_rhat,
            "max_abs_mean_err_over_sd": mean_err, "max_rel_var_err": var_err, "analysis_rows": view.shape[1],
            "valid": bool(valid)}


def bench_nfmcmc(args):
    import torch

    from flowmc_amd import random as fr
    from flowmc_amd import targets as TG
    from flowmc_amd.resource.logPDF import LogPDF
    from flowmc_amd.resource_strategy_bundle.RQSpline_MALA import RQSpline_MALA_Bundle
    from flowmc_amd.Sampler import Sampler

    dev = torch.device("cuda")
    C_, d = args.chains or 4096, 64
    cfg = nfmcmc_config(args)
    L, K, H = cfg["n_layers"], cfg["n_bins"], cfg["hidden"]
    target = TG.gaussian_mixture_64d(seed=0, device=dev)

    def make(n_chains, loops, key):
        k1, k2 = fr.split(key)
        b = RQSpline_MALA_Bundle(k1, n_chains, d, LogPDF(target, n_dims=d), cfg["n_local"], cfg["n_global"], loops,
                                 loops, cfg["n_epochs"], mala_step_size=cfg["step_size"], rq_spline_hidden_units=H,
                                 rq_spline_n_bins=K, rq_spline_n_layers=L, learning_rate=cfg["lr"],
                                 batch_size=cfg["batch_size"], n_max_examples=cfg["n_max_examples"],
                                 precision=args.precision, exact_inverse=cfg["exact_inverse"],
                                 history_window=cfg["history_window"])
        return b, Sampler(d, n_chains, k2, resource_strategy_bundles=b)

    x0 = nfmcmc_initial_positions(target, C_, dev, seed=1 + args.seed)
    # warm-up: a short run of the same shapes (kernels, allocator, workspaces)
    if args.warmup > 0:
        _, s = make(C_, args.warmup, fr.PRNGKey(7))
        s.sample(x0.clone(), {})
        torch.cuda.synchronize()
    bundle, sampler = make(C_, args.steps, fr.PRNGKey(args.seed))
    theta0 = bundle.resources["model"].theta.detach().clone()
    x_run = x0.clone()
    trace_marker()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sampler.sample(x_run, {})
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    trace_marker()
    res = sampler.resources
    an = nfmcmc_analysis(res["positions_production"].data, target, cfg["n_local"], cfg["n_global"])
    la = res["local_accs_production"].data
    ga = res["global_accs_production"].data
    local_acc = float(la[torch.isfinite(la)].mean())
    global_acc = float(ga[torch.isfinite(ga)].mean())
    value = an["min_ess_bulk"] / wall
    occupancy = mode_occupancy(res["positions_production"].data, target)
    # Rooflines of the three kernels that take ~90 % of the timed region's
    # kernel time, each timed alone with HIP events on its stream: the flow
    # adjoint of one layer at the training batch (algorithmic FLOPs: twice the
    # layer's forward over the batch), the layer's weight gradients (the
    # forward's multiply-adds, twice, over the batch) and the flow pass over
    # the proposals of a global step.
    import ctypes as _C

    import flowmc_amd._native as N

    model = res["model"]
    B = cfg["batch_size"]
    ws = model._workspace(B)
    model._ensure_packed()
    fs = model._struct()
    xb = res["positions_production"].data[:, -1].reshape(-1, d)
    xb = xb[torch.randint(0, xb.shape[0], (B,), device=dev)].contiguous()
    N.call("fmc_flow_pass", N.stream_ptr(), _C.byref(fs), 0, 1 | 2, xb.data_ptr(), B, 0, 0, 0, None,
           ws.logp.data_ptr(), ws.saved.data_ptr(), ws.g_logp.data_ptr(), ws.gy.data_ptr())
    lay = L - 1

    def adjoint():
        N.call("fmc_flow_backward_layer", N.stream_ptr(), _C.byref(fs), lay, ws.saved[lay].data_ptr(), B,
               ws.g_logp.data_ptr(), ws.gy.data_ptr(), ws.gx.data_ptr(), ws.store[0].data_ptr(), ws.ld,
               ws.aff[0].data_ptr(), ws.n_part)

    ms_adj = time_kernel(adjoint, reps=20)
    fl_adj = 2 * flow_layer_flops(d, lay, K, H) * B
    ach_adj = fl_adj / (ms_adj * 1e-3) / 1e12
    traffic_adj = None  # HBM bytes per launch from the counter runs of the flow_adjoint workload
    tfa = os.path.join(ROOT, "profiles", "flow_bwd_kernel_traffic.json")
    if os.path.exists(tfa) and B == 10000:
        with open(tfa) as f:
            traffic_adj = json.load(f).get("bytes_per_launch")
    # the weight gradients of the same layer (split-K product + fixed-order
    # sum of the partials): 2 * rows * cols FLOPs per sample and matrix
    adjoint()

    def wgrad():
        N.call("fmc_flow_weight_grad_split", N.stream_ptr(), _C.byref(fs), lay, ws.store[0].data_ptr(), ws.ld,
               ws.aff[0].data_ptr(), ws.n_part, ws.grad.data_ptr(), ws.wg_ws.data_ptr(), ws.wg_ws.numel() * 4)

    ms_wg = time_kernel(wgrad, reps=20)
    fl_wg = 2 * flow_wgrad_macs(d, lay, K, H) * B
    ach_wg = fl_wg / (ms_wg * 1e-3) / 1e12
    # the next kernel: one flow inverse pass over the C*S proposals of a global step
    S = cfg["n_global"]
    ms = time_kernel(lambda: model.sample(fr.PRNGKey(3), C_ * S), reps=3)
    fl = flow_flops_per_sample(d, L, K, H) * C_ * S
    achieved = fl / (ms * 1e-3) / 1e12
    traffic = None  # HBM bytes per launch from the counter runs of the same pass (flow_pass workload)
    tf = os.path.join(ROOT, "profiles", "flow_pass_kernel_traffic.json")
    if os.path.exists(tf) and C_ * S == 409600:
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    out = {
        "value": value,
        "ms_per_step": wall * 1e3 / (2 * args.steps),
        # the flow pass has the largest share of the timed region's kernel
        # time (600 training forwards at batch 10k, 40 sampling passes over
        # the 409,600 proposals of a global step, 40 log-densities); its
        # roofline is quoted at the sampling shape, the adjoint and the
        # weight gradients (the next two) beside it
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / BF16_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": "flow_pass_kernel (exact inverse over the 409,600 proposals of a global step, BF16)",
                     "kernel_ms": ms, "flops_per_launch": fl},
        "extra": {"wall_s": wall, **an, "local_acceptance": local_acc, "global_acceptance": global_acc,
                  "mode_occupancy": occupancy,
                  "roofline_adjoint": {"bound": "mfma", "achieved": ach_adj, "peak": BF16_PEAK_TFLOPS,
                                       "unit": "TFLOP/s", "frac": ach_adj / BF16_PEAK_TFLOPS, "traffic": traffic_adj,
                                       "kernel": f"flow_bwd_kernel_h256 (adjoint of one coupling layer, {B}-sample "
                                                 "batch, BF16)",
                                       "kernel_ms": ms_adj, "flops_per_launch": fl_adj},
                  "roofline_weight_grad": {"bound": "mfma", "achieved": ach_wg, "peak": BF16_PEAK_TFLOPS,
                                           "unit": "TFLOP/s", "frac": ach_wg / BF16_PEAK_TFLOPS,
                                           "kernel": "flow_wgrad_split_kernel + flow_wgrad_reduce_kernel (weight "
                                                     f"gradients of one coupling layer, {B}-sample batch)",
                                           "kernel_ms": ms_wg, "flops_per_launch": fl_wg}},
        "config": {"workload": "nfmcmc_gmm64", "chains": C_, "dim": d, "modes": 8, "mala_step": cfg["step_size"],
                   "flow": {"layers": L, "bins": K, "hidden": H, "precision": args.precision,
                            "exact_inverse": cfg["exact_inverse"]},
                   "training_rounds": args.steps, "production_rounds": args.steps, "local_steps": cfg["n_local"],
                   "global_steps": cfg["n_global"], "epochs": cfg["n_epochs"], "batch": cfg["batch_size"],
                   "n_max_examples": cfg["n_max_examples"], "lr": cfg["lr"],
                   "history_window": cfg["history_window"],
                   "initial_positions": "one exact mixture draw per chain"},
    }
    del res, sampler, bundle
    torch.cuda.empty_cache()
    if not args.no_cold:
        # The same run started from x0 ~ N(0, I): MALA at 0.05 reaches only the
        # basins near the start and the flow trains on what the chains hold,
        # so modes stay unvisited; reported for comparison (usually invalid).
        _, cs = make(C_, args.steps, fr.PRNGKey(args.seed))
        xc = torch.randn(C_, d, device=dev, generator=torch.Generator(device=dev).manual_seed(11 + args.seed))
        torch.cuda.synchronize()
        tc = time.perf_counter()
        cs.sample(xc, {})
        torch.cuda.synchronize()
        cwall = time.perf_counter() - tc
        cres = cs.resources
        can = nfmcmc_analysis(cres["positions_production"].data, target, cfg["n_local"], cfg["n_global"])
        cga = cres["global_accs_production"].data
        out["extra"]["cold_start"] = {
            "initial_positions": "x0 ~ N(0, I)", "value": can["min_ess_bulk"] / cwall, "wall_s": cwall, **can,
            "mode_occupancy": mode_occupancy(cres["positions_production"].data, target),
            "global_acceptance": float(cga[torch.isfinite(cga)].mean())}
        del cres, cs
        torch.cuda.empty_cache()
    if not args.no_plain:
        import bench_plain as BP

        print(f"bench.py: fused path {value:.0f} min-ESS/s in {wall:.2f} s; running the plain FP32 per-step path",
              file=sys.stderr, flush=True)

        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ppos, tr, pr = BP.run_plain_nfmcmc(theta0, cfg, target, x0.clone(), time_budget_s=args.plain_budget)
        torch.cuda.synchronize()
        pwall = time.perf_counter() - t0
        pan = nfmcmc_analysis(ppos, target, cfg["n_local"], cfg["n_global"])
        pval = pan["min_ess_bulk"] / pwall
        out["extra"]["plain_gpu"] = {"value": pval, "wall_s": pwall, **pan, "training_rounds": tr,
                                     "production_rounds": pr,
                                     "speedup_vs_plain": value / pval if pval > 0 else None}
        del ppos
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_nfmcmc(args, cfg, wall, an["min_ess_bulk"], C_)
    return out


def cpu_baseline_nfmcmc(args, cfg, gpu_wall, min_ess, C_):
    """The oracle (FP32 torch-CPU restatement of the reference: full-width
    conditioner, one-hot bins, 2 value_and_grad per MALA step) timed per
    component on bounded samples (one warm-up, median of 5, all host threads
    and one thread): MALA chain-steps, NF proposals (exact-inverse sample,
    its log_prob, the target) and training samples (forward + backward). The
    full run's CPU time is these rates times the run's counts; the fused
    run's min-ESS is divided by it (same algorithm, so the same ESS)."""
    import torch

    from oracle import chains as OC
    from oracle import flow as F
    from oracle import targets as OT
    from flowmc_amd import targets as TG

    d, L, K, H = 64, cfg["n_layers"], cfg["n_bins"], cfg["hidden"]
    tgt = TG.gaussian_mixture_64d(seed=0, device="cpu")
    fn = lambda z: OT.gmm_diag(z, tgt.means.float(), tgt.sigmas.float())
    p = F.init_params(d, L, K, H, dtype=torch.float32, hidden_var=1e-2)
    theta = F.to_flat(p).clone().requires_grad_(True)

    def mala(n):  # n chain-steps: up to 512 chains, 10 steps
        c = max(1, n // 10)
        x = torch.randn(c, d)

        def run():
            nonlocal x
            for _ in range(10):
                x, _, _, _ = OC.mala_step(x, fn, cfg["step_size"], torch.randn(c, d), torch.log(torch.rand(c)))
        return run

    def proposals(n):  # n global proposals: exact-inverse samples, their flow density and the target
        def run():
            with torch.no_grad():
                prop = F.sample_from_noise(p, torch.randn(n, d), K, exact=True)
                F.log_prob(p, prop, K)
                fn(prop)
        return run

    def train(n):  # n training samples: loss, forward + backward
        data = torch.randn(n, d)

        def run():
            loss = F.loss(F.from_flat(theta, d, L, K, H), data, K)
            torch.autograd.grad(loss, theta)
        return run

    r_mala = cpu_rate(mala, 5120, 640)
    r_prop = cpu_rate(proposals, 2048, 256)
    r_train = cpu_rate(train, 2000, 250)
    n_rounds = 2 * cfg["n_loops"]
    n_local = n_rounds * C_ * cfg["n_local"]
    n_global = n_rounds * C_ * cfg["n_global"]
    steps_per_round = cfg["n_epochs"] * max(cfg["n_max_examples"] // cfg["batch_size"], 1)
    n_train = cfg["n_loops"] * steps_per_round * cfg["batch_size"]
    tot = [n_local / r_mala[i] + n_global / r_prop[i] + n_train / r_train[i] for i in (0, 1)]
    return {"value": min_ess / tot[0], "unit": "min-ESS/s", "cores": r_mala[2], "kind": "port",
            "value_1_thread": min_ess / tot[1], "statistic": "median of 5 timed calls after one warm-up",
            "sample": f"FP32 torch-CPU oracle restatement ({cpu_model()}), {r_mala[2]} threads (1 thread on 1/8 of "
                      f"each sample): {r_mala[0]:.3g} MALA chain-steps/s (512 chains x 10 steps per call), "
                      f"{r_prop[0]:.3g} NF proposals/s (2048 per call), {r_train[0]:.3g} training samples/s "
                      f"(2000 per call); the run's {n_local:.3g} chain-steps, {n_global:.3g} proposals and "
                      f"{n_train:.3g} training samples at those rates = {tot[0]:.0f} s ({tot[1]:.0f} s on 1 "
                      f"thread); min-ESS of the fused run (same algorithm)"}


# --------------------------------------------------------------- C3 -------
def bench_mala_rosenbrock(args):
    return _bench_local_rosenbrock(args, "mala")


def bench_grw_rosenbrock(args):
    """Fused Gaussian random walk on the same 128-D Rosenbrock target (16384
    chains, step 0.01, 64 steps per launch)."""
    return _bench_local_rosenbrock(args, "grw")


def _bench_local_rosenbrock(args, kind):
    import torch

    import flowmc_amd._native as N
    from flowmc_amd import random as fr
    from flowmc_amd import targets as TG

    C_, d, dt, K = args.chains or 16384, 128, 0.01, 64
    dev = torch.device("cuda")
    tgt = TG.Rosenbrock(d, device=dev)
    desc = tgt.descriptor()
    steps = args.steps
    T = max(steps, args.warmup)
    x = fr.normal(fr.PRNGKey(0), (C_, d), device=dev)
    lp = torch.empty(C_, dtype=torch.float32, device=dev)
    pos = torch.empty((C_, T, d), dtype=torch.float32, device=dev)
    lps = torch.empty((C_, T), dtype=torch.float32, device=dev)
    acc = torch.empty((C_, T), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream()
    cfg = N.fmc_launch_cfg(args.lpc, args.nv, args.block, 0) if args.lpc else None
    cfgp = C.byref(cfg) if cfg is not None else None

    def run(n_steps, key):
        k0, k1 = fr.key_words(key)
        st = torch.cuda.current_stream().cuda_stream  # the capture stream while a graph is recorded
        for s0 in range(0, n_steps, K):
            ns = min(K, n_steps - s0)
            N.call("fmc_mala_run" if kind == "mala" else "fmc_grw_run", st, C.byref(desc), x.data_ptr(),
                   lp.data_ptr(), C_, ns, s0, dt, k0, k1, None, None, pos.data_ptr(), lps.data_ptr(), acc.data_ptr(),
                   T, 0, 1, cfgp)

    # calibration copy of known traffic (256 MiB read + 256 MiB written) for
    # the FETCH_SIZE / WRITE_SIZE counter runs of this workload
    ca = torch.empty(64 << 20, dtype=torch.float32, device=dev)
    cb = torch.empty_like(ca)
    cb.copy_(ca)
    del ca, cb
    run(args.warmup, fr.PRNGKey(1))
    torch.cuda.synchronize()
    graph = None
    if not args.no_graph:
        # the whole timed sequence of fused launches (64 steps each) as one HIP graph
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            run(steps, fr.PRNGKey(2))
        torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    if graph is not None:
        graph.replay()
    else:
        run(steps, fr.PRNGKey(2))
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    n_launch = (steps + K - 1) // K
    bytes_total = C_ * (steps * (4 * d + 8) + n_launch * (8 * d + 8))
    achieved = bytes_total / (gpu_ms * 1e-3) / 1e9
    flops = C_ * steps * ((17 * d + 12 * d) if kind == "mala" else (8 * d + 8 * d))
    lay = f"{args.lpc} lanes x {args.nv} float4 per chain" if args.lpc else "default layout: 16 lanes x 2 float4 per chain"
    kname = (f"chain_kernel<{args.nv if args.lpc else 2},{kind.upper()},ROSENBROCK> "
             f"(fused {kind.upper()}, {lay})")
    traffic = None  # HBM bytes per 64-step launch from the counter runs of this workload
    tf = os.path.join(ROOT, "profiles", f"chain_kernel_traffic__{kind}_rosenbrock_128d.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    out = {
        "value": C_ * steps / wall,
        "ms_per_step": wall * 1e3 / steps,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kname,
                     "kernel_ms": gpu_ms / n_launch, "bytes_per_launch": bytes_total / n_launch},
        "extra": {"fp32_tflops": flops / (gpu_ms * 1e-3) / 1e12, "acceptance": float(acc[:, :steps].mean())},
        "config": {"workload": f"{kind}_rosenbrock_128d", "chains": C_, "dim": d, "step_size": dt,
                   "fused_steps_per_launch": K, "thinning": 1,
                   "launch": "eager" if graph is None else f"one HIP graph of {n_launch} fused launches"},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_mala_rosenbrock(args, kind)
    return out


def cpu_baseline_mala_rosenbrock(args, kind="mala"):
    """Oracle MALA (2 value_and_grad per step as MALA.py:59,71-73) or GRW (one
    target evaluation per step), FP32, timed with cpu_rate on bounded samples
    (1024 chains x 8 steps per call; one thread: 128 chains)."""
    import torch

    from oracle import chains as OC
    from oracle import targets as OT

    d, dt = 128, 0.01
    fn = lambda z: OT.rosenbrock(z)

    def work(n):
        c = max(1, n // 8)
        x = torch.randn(c, d)
        lp = fn(x)

        def run():
            nonlocal x, lp
            for _ in range(8):
                xi, lu = torch.randn(c, d), torch.log(torch.rand(c))
                if kind == "mala":
                    x, _, _, _ = OC.mala_step(x, fn, dt, xi, lu)
                else:
                    x, lp, _, _ = OC.grw_step(x, lp, fn, dt, xi, lu)
        return run

    r = cpu_rate(work, 8192, 1024)
    return {"value": r[0], "unit": "chain-steps/s", "cores": r[2], "kind": "port", "value_1_thread": r[1],
            "statistic": "median of 5 timed calls after one warm-up",
            "sample": f"1024 chains x 8 {kind.upper()} steps per call (1 thread: 128 chains), d={d}, FP32 "
                      f"torch-CPU oracle restatement ({cpu_model()})"}


# --------------------------------------------------------------- C2 -------
def bench_flow_train(args):
    import torch

    from flowmc_amd import random as fr
    from flowmc_amd.resource.model.nf_model.rqSpline import MaskedCouplingRQSpline
    from flowmc_amd.resource.optimizer import adam

    dev = torch.device("cuda")
    d, L, K, H, N, B = 32, 8, 10, [128, 128], 100000, 10000
    g = torch.Generator().manual_seed(0)
    means = torch.rand(4, d, generator=g) * 10 - 5
    lab = torch.randint(0, 4, (N,), generator=g)
    data = (means[lab] + torch.randn(N, d, generator=g)).to(dev)
    m = MaskedCouplingRQSpline(d, L, H, K, fr.PRNGKey(0), precision=args.precision)
    opt = adam(1e-3)
    state = opt.init(m)
    m.set_data_stats(data.mean(0), torch.cov(data.T))
    perm = torch.randperm(N, device=dev)
    batches = [data[perm[i * B:(i + 1) * B]].contiguous() for i in range(N // B)]
    for i in range(max(1, args.warmup)):
        m.train_step(batches[i % len(batches)], opt, state)
    torch.cuda.synchronize()
    l0 = float(-m.log_prob(data[:20000]).mean())
    torch.cuda.synchronize()
    trace_marker()
    t0 = time.perf_counter()
    for i in range(args.steps):
        loss, m, state = m.train_step(batches[i % len(batches)], opt, state)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    trace_marker()
    l1 = float(-m.log_prob(data[:20000]).mean())
    # dominant kernel: the per-layer adjoint + weight-gradient; report the whole step on MFMA
    fl = 3 * flow_flops_per_sample(d, L, K, H) * B
    achieved = fl * args.steps / wall / 1e12
    ms_fwd = time_kernel(lambda: m.log_prob(batches[0]), reps=5)
    out = {
        "value": args.steps * B / wall,
        "ms_per_step": wall * 1e3 / args.steps,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / BF16_PEAK_TFLOPS, "traffic": None,
                     "kernel": "train step (forward + adjoint + weight grad + AdamW)",
                     "forward_pass_ms": ms_fwd,
                     "forward_pass_tflops": flow_flops_per_sample(d, L, K, H) * B / (ms_fwd * 1e-3) / 1e12},
        "extra": {"loss_before": l0, "loss_after": l1},
        "config": {"workload": "flow_train_d32", "dim": d, "layers": L, "bins": K, "hidden": H, "batch": B,
                   "examples": N, "optimizer": "adam(1e-3)", "precision": args.precision},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_flow_train(args, d, L, K, H, B)
    return out


def cpu_baseline_flow_train(args, d, L, K, H, B):
    """Oracle loss + autograd gradient (FP32 torch-CPU) per training sample,
    timed with cpu_rate (2000-sample batches; one thread: 250)."""
    import torch

    from oracle import flow as F

    p = F.init_params(d, L, K, H, dtype=torch.float32, hidden_var=1e-2)
    theta = F.to_flat(p).clone().requires_grad_(True)

    def work(n):
        x = torch.randn(n, d)

        def run():
            loss = F.loss(F.from_flat(theta, d, L, K, H), x, K)
            torch.autograd.grad(loss, theta)
        return run

    r = cpu_rate(work, 2000, 250)
    return {"value": r[0], "unit": "samples/s", "cores": r[2], "kind": "port", "value_1_thread": r[1],
            "statistic": "median of 5 timed calls after one warm-up",
            "sample": f"forward + backward of 2000-sample batches (1 thread: 250), FP32 torch-CPU oracle "
                      f"({cpu_model()})"}


# ------------------------------------------------- flow pass (profiling) --
def bench_flow_pass(args):
    """The dominant kernel of the NF-MCMC workload alone: one exact-inverse
    flow pass over the 409,600 proposals of a global step (d=64, 10 layers, 8
    bins, [256, 256]), ``--steps`` launches. It also runs one 256 MiB
    device-to-device copy (a kernel of known traffic) so that counter runs
    (FETCH_SIZE / WRITE_SIZE) can be scaled to bytes. Traffic measured that
    way is read from profiles/flow_pass_kernel_traffic.json when present."""
    import torch

    from flowmc_amd import random as fr
    from flowmc_amd.resource.model.nf_model.rqSpline import MaskedCouplingRQSpline

    dev = torch.device("cuda")
    d, L, K, H, n = 64, 10, 8, [256, 256], 409600
    m = MaskedCouplingRQSpline(d, L, H, K, fr.PRNGKey(0), precision=args.precision, device=dev, exact_inverse=True)
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)  # calibration copy: 256 MiB read + 256 MiB written
    for i in range(max(1, args.warmup)):
        m.sample(fr.PRNGKey(i), n)
    ms = time_kernel(lambda: m.sample(fr.PRNGKey(1), n), reps=args.steps)
    fl = flow_flops_per_sample(d, L, K, H) * n
    achieved = fl / (ms * 1e-3) / 1e12
    traffic = None
    tf = os.path.join(ROOT, "profiles", "flow_pass_kernel_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    return {
        "value": n / (ms * 1e-3),
        "ms_per_step": ms,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS if args.precision == "bfloat16"
                     else FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / (BF16_PEAK_TFLOPS if args.precision == "bfloat16" else FP32_PEAK_TFLOPS),
                     "traffic": traffic, "kernel": "flow_pass_kernel (exact inverse)", "kernel_ms": ms,
                     "flops_per_launch": fl, "algorithmic_bytes_per_launch": n * (d * 4 * 2)},
        "config": {"workload": "flow_pass", "samples": n, "dim": d, "layers": L, "bins": K, "hidden": H,
                   "precision": args.precision},
    }


# ---------------------------------------------------------------- RealNVP --
def bench_realnvp_train(args):
    """RealNVP training (d=32, 8 layers, hidden 256, batch 10k, AdamW with
    clipping) on a synthetic 4-mode mixture; a step is one optimizer step. The
    sampling pass (409,600 samples from base_dist) is timed beside it."""
    import torch

    from flowmc_amd import random as fr
    from flowmc_amd.resource.model.nf_model.realNVP import RealNVP
    from flowmc_amd.resource.optimizer import adamw

    dev = torch.device("cuda")
    d, L, H, N, B = 32, 8, 256, 100000, 10000
    g = torch.Generator().manual_seed(0)
    means = torch.rand(4, d, generator=g) * 6 - 3
    lab = torch.randint(0, 4, (N,), generator=g)
    data = (means[lab] + torch.randn(N, d, generator=g)).to(dev)
    m = RealNVP(d, L, H, fr.PRNGKey(0), device=dev)
    m.set_data_stats(data.mean(0), torch.cov(data.T))
    opt = adamw(1e-3, max_norm=1.0)
    state = opt.init(m)
    perm = torch.randperm(N, device=dev)
    batches = [data[perm[i * B:(i + 1) * B]].contiguous() for i in range(N // B)]
    for i in range(max(1, args.warmup)):
        m.train_step(batches[i % len(batches)], opt, state)
    torch.cuda.synchronize()
    l0 = float(-m.log_prob(data[:20000]).mean())
    torch.cuda.synchronize()
    trace_marker()
    t0 = time.perf_counter()
    for i in range(args.steps):
        loss, m, state = m.train_step(batches[i % len(batches)], opt, state)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    trace_marker()
    l1 = float(-m.log_prob(data[:20000]).mean())
    fl_fwd = 4 * H * d * L  # two MLPs, d_c + d_t = d inputs + outputs, per sample
    fl_step = 4 * fl_fwd * B  # forward + adjoint (recompute, two transposed products, two weight products)
    achieved = fl_step * args.steps / wall / 1e12
    n_s = 409600
    ms_s = time_kernel(lambda: m.sample(fr.PRNGKey(1), n_s), reps=5)
    out = {
        "value": args.steps * B / wall,
        "ms_per_step": wall * 1e3 / args.steps,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": "train step (fused forward + per-layer adjoint + partial reduction + AdamW), FP32",
                     "flops_per_step": fl_step},
        "extra": {"loss_before": l0, "loss_after": l1, "sample_ms_409600": ms_s,
                  "sample_tflops": fl_fwd * n_s / (ms_s * 1e-3) / 1e12},
        "config": {"workload": "realnvp_train_d32", "dim": d, "layers": L, "hidden": H, "batch": B, "examples": N,
                   "optimizer": "adamw(1e-3) + clip 1.0"},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_realnvp(args, d, L, H, B)
    return out


def cpu_baseline_realnvp(args, d, L, H, B):
    """Oracle RealNVP loss + autograd gradient (FP32 torch-CPU) per training
    sample, timed with cpu_rate (2000-sample batches; one thread: 250)."""
    import torch

    from oracle import realnvp as R

    p = R.init_params(d, L, H, dtype=torch.float32)
    theta = R.to_flat(p).clone().requires_grad_(True)

    def work(n):
        x = torch.randn(n, d)

        def run():
            loss = R.loss(R.from_flat(theta, d, L, H), x)
            torch.autograd.grad(loss, theta)
        return run

    r = cpu_rate(work, 2000, 250)
    return {"value": r[0], "unit": "samples/s", "cores": r[2], "kind": "port", "value_1_thread": r[1],
            "statistic": "median of 5 timed calls after one warm-up",
            "sample": f"forward + backward of 2000-sample batches (1 thread: 250), FP32 torch-CPU oracle "
                      f"({cpu_model()})"}


# ----------------------------------------------------- parallel tempering --
def bench_pt_gmm64(args):
    """ParallelTempering in training mode on the 64-D 8-mode mixture: 4096
    chains x 5 temperatures (linspace(1, 5), the bundle's ladder), 100 tempered
    MALA steps (step 0.05) per call, then the swap sweep and the temperature
    adaptation. A step is one strategy call; the value counts tempered
    chain-steps (chains x temperatures x MALA steps) per second of wall time.
    The dominant kernel is the fused tempered MALA launch (one per call),
    timed with HIP events on the launch stream in a separate pass."""
    import torch

    from flowmc_amd import random as fr
    from flowmc_amd import targets as TG
    from flowmc_amd.resource.buffers import Buffer
    from flowmc_amd.resource.kernel.MALA import MALA
    from flowmc_amd.resource.logPDF import LogPDF, TemperedPDF
    from flowmc_amd.resource.states import State
    from flowmc_amd.strategy.parallel_tempering import ParallelTempering

    dev = torch.device("cuda")
    C_, T, d, S, dt = args.chains or 4096, 5, 64, 100, 0.05
    tgt = TG.gaussian_mixture_64d(seed=0, device=dev)
    tpdf = TemperedPDF(LogPDF(tgt, n_dims=d), None, n_dims=d, n_temps=T)
    temps = Buffer("temperature", (T,), 0, device=dev)
    temps.update_buffer(torch.linspace(1.0, 5.0, T, dtype=torch.float64).float())
    tpos = Buffer("tempered_positions", (C_, T - 1, d), 2, device=dev)
    x = nfmcmc_initial_positions(tgt, C_, dev)
    tpos.update_buffer(x[:, None].expand(C_, T - 1, d))
    import flowmc_amd._native as N

    cfg = N.fmc_launch_cfg(args.lpc, args.nv, args.block, 0) if args.lpc else None
    res = {"tlp": tpdf, "mala": MALA(dt, fused_steps=S, launch_cfg=cfg), "tp": tpos, "T": temps,
           "state": State({"training": True}, name="sampler_state")}
    strat = ParallelTempering(S, "tlp", "mala", ["tp", "T"], "state")
    key = fr.PRNGKey(args.seed)
    for _ in range(args.warmup):
        key, res, x = strat(key, res, x, {})
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        key, res, x = strat(key, res, x, {})
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # the dominant kernel alone: the fused tempered MALA launch over C*T rows
    stream = torch.cuda.current_stream()
    pos = torch.cat([x[:, None], tpos.data], 1).contiguous()
    acc = torch.empty(C_ * T, S, dtype=torch.float32, device=dev)
    reps = 5
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for i in range(reps):
        pos, _ = res["mala"].sample_tempered(fr.PRNGKey(100 + i), pos, tpdf, temps.data, None, S, accepts=acc)
    ev1.record(stream)
    torch.cuda.synchronize()
    k_ms = ev0.elapsed_time(ev1) / reps
    rows, K = C_ * T, tgt.n_modes
    bytes_launch = rows * (8 * d + 4 + 4 * S)  # state in + out, log-density out, decisions
    achieved = bytes_launch / (k_ms * 1e-3) / 1e9
    flops = rows * S * (7 * d * K + 10 * d)  # GMM value + gradient per step, MALA update and ratio
    t_final = res["T"].data.tolist()
    return {
        "value": C_ * T * S * args.steps / wall,
        "ms_per_step": wall * 1e3 / args.steps,
        # compute-bound: the 8-mode value + gradient per step; the HBM side is in extra
        "roofline": {"bound": "valu", "achieved": flops / (k_ms * 1e-3) / 1e12, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": flops / (k_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": "chain_kernel (fused tempered MALA, FP32 vector ALU, one launch of 100 steps per call)",
                     "kernel_ms": k_ms, "flops_per_launch": flops},
        "extra": {"hbm_gbs": achieved, "hbm_frac": achieved / HBM_PEAK_GBS, "bytes_per_launch": bytes_launch,
                  "local_acceptance": float(acc.mean()), "temperatures_after": t_final},
        "config": {"workload": "pt_mala_gmm64", "chains": C_, "temperatures": T, "max_temperature": 5.0,
                   "tempered_steps": S, "step_size": dt, "dim": d, "modes": K,
                   "layout": [args.lpc, args.nv, args.block] if args.lpc else "default"},
    } | ({} if args.no_cpu_baseline else {"cpu_baseline": cpu_baseline_pt(args, tgt, T, S, dt)})


def cpu_baseline_pt(args, tgt, T, S, dt):
    """Oracle parallel tempering (tempered MALA steps, then the swap sweep and
    the adaptation), FP32 torch on the host cores, timed with cpu_rate: one
    call = 10 tempered MALA steps of 128 chains x T temperatures and one
    exchange + adaptation (one thread: 16 chains)."""
    import torch

    from oracle import chains as OC
    from oracle import targets as OT
    from oracle import tempering as OPT

    d = tgt.n_dim
    means, sigmas = tgt.means.float().cpu(), tgt.sigmas.float().cpu()
    fn = lambda z: OT.gmm_diag(z, means, sigmas)
    n_steps = 10

    def work(n):  # n tempered chain-steps
        c = max(1, n // (T * n_steps))
        temps = torch.linspace(1.0, 5.0, T)
        x = means[torch.randint(0, means.shape[0], (c * T,))] + torch.randn(c * T, d)

        def run():
            nonlocal x, temps
            fr_ = OPT.tempered_rows(fn, temps, c)
            for _ in range(n_steps):
                x, _, _, _ = OC.mala_step(x, lambda z: fr_(z).to(z.dtype), dt, torch.randn(c * T, d),
                                          torch.log(torch.rand(c * T)))
            xs, _, acc = OPT.exchange(x.reshape(c, T, d), fn(x).reshape(c, T), temps,
                                      torch.log(torch.rand(c, T - 1)))
            temps = OPT.adapt_temperature(temps, acc).float()
            x = xs.reshape(c * T, d)
        return run

    r = cpu_rate(work, 128 * T * n_steps, 16 * T * n_steps)
    return {"value": r[0], "unit": "tempered chain-steps/s", "cores": r[2], "kind": "port", "value_1_thread": r[1],
            "statistic": "median of 5 timed calls after one warm-up",
            "sample": f"128 chains x {T} temperatures x {n_steps} tempered MALA steps + one swap sweep and "
                      f"adaptation per call (1 thread: 16 chains), FP32 torch-CPU oracle restatement ({cpu_model()})"}


# --------------------------------------------------------------- C4 -------
def bench_hmc_funnels(args):
    """Batched uneven-dimension HMC: 32 Neal's funnels, d log-uniform in
    [8, 256], 256 chains each, 10 leapfrog steps of 0.05, 16 steps per launch.

    The step is FP32-vector bound, not HBM bound: per chain-step it computes
    (L+2) gradients (3d each), (L+1) drift+kick pairs (4d) and the kinetic
    terms (6d) = 86d FLOPs for L = 10, against 4d+8 bytes of row stores. The
    roofline is that FLOP count over the HIP-event time of the launches on
    the launching stream; the HBM figure of the same launches is in extra."""
    import torch

    from flowmc_amd import random as fr
    from flowmc_amd import targets as TG
    from flowmc_amd.resource.kernel.batched_HMC import BatchedHMC

    dims = TG.funnel_dims(seed=0, n_targets=32, lo=8, hi=256)
    C_, L, K = args.chains or 256, 10, 16
    bh = BatchedHMC([TG.Funnel(d) for d in dims], C_, step_size=0.05, n_leapfrog=L)
    bh.init_positions(fr.PRNGKey(0))
    T = args.steps
    outs = [(torch.empty((C_, T, d), dtype=torch.float32, device="cuda"),
             torch.empty((C_, T), dtype=torch.float32, device="cuda"),
             torch.empty((C_, T), dtype=torch.float32, device="cuda"), 0) for d in dims]
    layout = bh.tune(fr.PRNGKey(9))
    # calibration copy of known traffic (256 MiB read + 256 MiB written) for
    # the FETCH_SIZE / WRITE_SIZE counter runs of this workload
    ca = torch.empty(64 << 20, dtype=torch.float32, device="cuda")
    cb = torch.empty_like(ca)
    cb.copy_(ca)
    del ca, cb
    bh.run(fr.PRNGKey(1), min(T, K), outputs=outs)
    torch.cuda.synchronize()
    launch = bh.prepare(fr.PRNGKey(2), T, outputs=outs, fused_steps=K)  # tables uploaded here
    graph = None
    if not args.no_graph:
        # the timed launches (16 steps each, all 32 targets per launch) as one HIP graph
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            launch()
        torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    if graph is not None:
        graph.replay()
    else:
        launch()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = ev0.elapsed_time(ev1) * 1e-3
    n_launch = (T + K - 1) // K
    sum_d = sum(dims)
    flops_total = (L + 2) * 3 * sum_d + 4 * (L + 1) * sum_d + 6 * sum_d
    flops_total *= C_ * T
    bytes_total = sum(C_ * T * (4 * d + 8) + C_ * n_launch * (8 * d + 8) for d in dims)
    achieved = flops_total / gpu_s / 1e12
    traffic = None  # HBM bytes per 16-step launch from the counter runs of this workload
    tf = os.path.join(ROOT, "profiles", "hmc_batched_unit_kernel_traffic__hmc_funnels.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    v_all = torch.cat([o[0][..., 0].reshape(-1) for o in outs])
    out = {
        "value": C_ * len(dims) * T / wall,
        "ms_per_step": wall * 1e3 / T,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": ("hmc_batched_unit_kernel" if bh.uniform_kind() >= 0 else "hmc_batched_kernel")
                     + " (FP32 vector ALU)", "kernel_ms": gpu_s * 1e3 / n_launch,
                     "flops_per_launch": flops_total / n_launch},
        "extra": {"hbm_gbs": bytes_total / gpu_s / 1e9, "hbm_frac": bytes_total / gpu_s / 1e9 / HBM_PEAK_GBS,
                  "acceptance": float(torch.cat([o[2].reshape(-1) for o in outs]).mean()),
                  "v_mean": float(v_all.mean()), "v_var": float(v_all.var()), "layout": [layout[0], layout[1]],
                  "dims": dims},
        "config": {"workload": "hmc_funnels", "targets": 32, "chains_per_target": C_, "n_leapfrog": L,
                   "step_size": 0.05, "fused_steps_per_launch": K,
                   "launch": "eager" if graph is None else f"one HIP graph of {n_launch} batched launches"},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_hmc_funnels(args, dims, C_, L)
    return out


def cpu_baseline_hmc_funnels(args, dims, C_, L):
    """Oracle HMC steps (autograd gradients, FP32 torch on the host cores):
    one call = one HMC step of every funnel of the batch with 64 chains each
    (one thread: 8), timed with cpu_rate; the rate counts chain-steps."""
    import torch

    from oracle import chains as OC
    from oracle import targets as OT

    def work(n):  # n chain-steps over all funnels
        c = max(1, n // len(dims))
        xs = [torch.randn(c, d) for d in dims]
        lps = [OT.funnel(x) for x in xs]
        ones = [torch.ones(d) for d in dims]

        def run():
            for i, d in enumerate(dims):
                xs[i], lps[i], _, _ = OC.hmc_step(xs[i], lps[i], OT.funnel, 0.05, L, ones[i], torch.randn(c, d),
                                                  torch.log(torch.rand(c)))
        return run

    r = cpu_rate(work, 64 * len(dims), 8 * len(dims))
    return {"value": r[0], "unit": "chain-steps/s", "cores": r[2], "kind": "port", "value_1_thread": r[1],
            "statistic": "median of 5 timed calls after one warm-up",
            "sample": f"one HMC step (L={L}) of all {len(dims)} funnels x 64 chains per call (1 thread: 8 chains), "
                      f"FP32 torch-CPU oracle restatement with autograd gradients ({cpu_model()})"}


# --------------------------------------------------------------- C1 -------
def bench_mala_gauss5d(args):
    """MALA on the 5-D correlated Gaussian (Sigma = Q diag(lam) Q^T, Q seeded
    orthogonal, lam log-uniform in [0.1, 10]): 1000 chains from x0 ~ N(0, I),
    step 0.1, ``--steps`` (1000) steps through TakeSerialSteps, all steps in
    one fused launch, then one TrainModel round on the chain buffer
    (MaskedCouplingRQSpline 4 layers / 8 bins / [32, 32], clip + AdamW lr
    1e-3, batch 1000, 5 epochs over n_max_examples 10000 from the last 100
    rows of every chain). value = chain-steps/s of the MALA phase; the
    training round is reported beside it. At 1000 chains the fused kernel
    occupies a few dozen wavefronts: it is latency-bound, far below the HBM
    roofline its bytes would allow."""
    import torch

    from flowmc_amd import random as fr
    from flowmc_amd import targets as TG
    from flowmc_amd.resource.buffers import Buffer
    from flowmc_amd.resource.kernel.MALA import MALA
    from flowmc_amd.resource.logPDF import LogPDF
    from flowmc_amd.resource.model.nf_model.rqSpline import MaskedCouplingRQSpline
    from flowmc_amd.resource.optimizer import Optimizer
    from flowmc_amd.resource.states import State
    from flowmc_amd.strategy.take_steps import TakeSerialSteps
    from flowmc_amd.strategy.train_model import TrainModel

    dev = torch.device("cuda")
    C_, d, dt, S = args.chains or 1000, 5, 0.1, args.steps
    tgt = TG.correlated_gaussian_5d(seed=0, device=dev)

    def resources(n_steps):
        return {"positions": Buffer("positions", (C_, n_steps, d), 1, device=dev),
                "log_prob": Buffer("log_prob", (C_, n_steps), 1, device=dev),
                "acceptance": Buffer("acceptance", (C_, n_steps), 1, device=dev),
                "mala": MALA(dt, fused_steps=max(n_steps, 1)), "logpdf": LogPDF(tgt, n_dims=d),
                "sampler_state": State({"positions": "positions", "log_prob": "log_prob",
                                        "acceptance": "acceptance"}, name="sampler_state")}

    def strategy(n_steps):
        return TakeSerialSteps("logpdf", kernel_name="mala", state_name="sampler_state",
                               buffer_names=["positions", "log_prob", "acceptance"], n_steps=n_steps)

    x0 = torch.randn(C_, d, device=dev, generator=torch.Generator(device=dev).manual_seed(args.seed))
    if args.warmup > 0:
        strategy(args.warmup)(fr.PRNGKey(5), resources(args.warmup), x0.clone(), {})
    res = resources(S)
    strat = strategy(S)
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    trace_marker()
    t0 = time.perf_counter()
    ev0.record(stream)
    _, res, last = strat(fr.PRNGKey(args.seed), res, x0.clone(), {})
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    trace_marker()
    gpu_ms = ev0.elapsed_time(ev1)
    pos = res["positions"].data
    acc = res["acceptance"].data
    # the TrainModel round on the chain buffer
    model = MaskedCouplingRQSpline(d, 4, [32, 32], 8, fr.PRNGKey(1), precision=args.precision)
    res["model"] = model
    res["optimizer"] = Optimizer(model, learning_rate=1e-3)
    res["loss"] = Buffer("loss", (5,), 0, device=dev)
    train = TrainModel("model", "positions", "optimizer", loss_buffer_name="loss", n_epochs=5, batch_size=1000,
                       n_max_examples=10000, history_window=100)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _, res, _ = train(fr.PRNGKey(2), res, last, {})
    torch.cuda.synchronize()
    train_wall = time.perf_counter() - t1
    losses = res["loss"].data.tolist()
    cov_emp = torch.cov(pos[:, S // 2:].reshape(-1, d).T.double())
    cov_t = tgt.covariance.to(dev)
    bytes_total = C_ * (S * (4 * d + 8) + (8 * d + 8))
    achieved = bytes_total / (gpu_ms * 1e-3) / 1e9
    out = {
        "value": C_ * S / wall,
        "ms_per_step": wall * 1e3 / S,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": f"chain_kernel<MALA, GAUSS_DENSE> ({S} steps in one launch; latency-bound at "
                               f"{C_} chains)", "kernel_ms": gpu_ms, "bytes_per_launch": bytes_total,
                     "ns_per_step": gpu_ms * 1e6 / S},
        "extra": {"acceptance": float(acc.mean()), "train_wall_s": train_wall,
                  "train_samples_per_s": 5 * 10000 / train_wall, "train_losses": losses,
                  "c1_wall_s": wall + train_wall,
                  "max_rel_cov_err_second_half": float(((cov_emp - cov_t).abs() / cov_t.diagonal().sqrt()[:, None]
                                                        / cov_t.diagonal().sqrt()[None, :]).max())},
        "config": {"workload": "mala_gauss5d", "chains": C_, "dim": d, "step_size": dt, "steps": S,
                   "initial_positions": "x0 ~ N(0, I)",
                   "train": {"layers": 4, "bins": 8, "hidden": [32, 32], "optimizer": "clip(1) + adamw(1e-3)",
                             "batch": 1000, "epochs": 5, "n_max_examples": 10000, "history_window": 100,
                             "precision": args.precision}},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_gauss5d(tgt, C_, d, dt, S)
    return out


def cpu_baseline_gauss5d(tgt, C_, d, dt, S):
    """Oracle MALA on the same 5-D Gaussian (2 value_and_grad per step) and the
    oracle flow's training step, FP32 torch-CPU, timed with cpu_rate: 1000
    chains x 10 steps per call (one thread: 125 chains); 1000-sample batches
    (one thread: 125). The whole workload's CPU time is extrapolated from the
    two rates in extra."""
    import torch

    from oracle import chains as OC
    from oracle import flow as F
    from oracle import targets as OT

    P = tgt.precision.double().cpu().float()
    fn = lambda z: OT.gauss_dense(z, torch.zeros(d), P)

    def mala(n):
        c = max(1, n // 10)
        x = torch.randn(c, d)

        def run():
            nonlocal x
            for _ in range(10):
                x, _, _, _ = OC.mala_step(x, fn, dt, torch.randn(c, d), torch.log(torch.rand(c)))
        return run

    p = F.init_params(d, 4, 8, [32, 32], dtype=torch.float32, hidden_var=1e-2)
    theta = F.to_flat(p).clone().requires_grad_(True)

    def train(n):
        x = torch.randn(n, d)

        def run():
            torch.autograd.grad(F.loss(F.from_flat(theta, d, 4, 8, [32, 32]), x, 8), theta)
        return run

    r = cpu_rate(mala, C_ * 10, C_ * 10 // 8)
    rt = cpu_rate(train, 1000, 125)
    total = C_ * S / r[0] + 50000 / rt[0]
    return {"value": r[0], "unit": "chain-steps/s", "cores": r[2], "kind": "port", "value_1_thread": r[1],
            "statistic": "median of 5 timed calls after one warm-up",
            "train_samples_per_s": rt[0], "c1_cpu_s": total,
            "sample": f"{C_} chains x 10 MALA steps per call (1 thread: {C_ // 8} chains), d={d}; training: "
                      f"1000-sample forward + backward per call; FP32 torch-CPU oracle restatement ({cpu_model()}); "
                      f"the workload's {C_ * S} chain-steps and 50,000 training samples take {total:.1f} s at "
                      f"those rates"}


# --------------------------------------------------- flow adjoint (profiling) --
def bench_flow_adjoint(args):
    """The dominant kernel of the NF-MCMC workload's timed region alone: the
    adjoint of one coupling layer (d=64, [256, 256], 8 bins, BF16) over a
    10,000-sample training batch, ``--steps`` launches. Also runs one 256 MiB
    device-to-device copy (a kernel of known traffic) so that counter runs
    (FETCH_SIZE / WRITE_SIZE) can be scaled to bytes; traffic measured that
    way is read from profiles/flow_bwd_kernel_traffic.json when present."""
    import ctypes as _C

    import torch

    import flowmc_amd._native as N
    from flowmc_amd import random as fr
    from flowmc_amd.resource.model.nf_model.rqSpline import MaskedCouplingRQSpline

    dev = torch.device("cuda")
    d, L, K, H, B = 64, 10, 8, [256, 256], 10000
    m = MaskedCouplingRQSpline(d, L, H, K, fr.PRNGKey(0), precision=args.precision, device=dev)
    x = torch.randn(B, d, device=dev) * 2
    m.grad_log_prob_sum(x)  # builds the workspace and the saved layer inputs
    ws = m._workspace(B)
    fs = m._struct()
    N.call("fmc_flow_pass", N.stream_ptr(), _C.byref(fs), 0, 1 | 2, x.data_ptr(), B, 0, 0, 0, None,
           ws.logp.data_ptr(), ws.saved.data_ptr(), ws.g_logp.data_ptr(), ws.gy.data_ptr())
    a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)  # calibration copy: 256 MiB read + 256 MiB written
    lay = L - 1

    def adjoint():
        N.call("fmc_flow_backward_layer", N.stream_ptr(), _C.byref(fs), lay, ws.saved[lay].data_ptr(), B,
               ws.g_logp.data_ptr(), ws.gy.data_ptr(), ws.gx.data_ptr(), ws.store[0].data_ptr(), ws.ld,
               ws.aff[0].data_ptr(), ws.n_part)

    for _ in range(max(1, args.warmup)):
        adjoint()
    trace_marker()
    ms = time_kernel(adjoint, reps=args.steps)
    trace_marker()
    fl = 2 * flow_layer_flops(d, lay, K, H) * B
    achieved = fl / (ms * 1e-3) / 1e12
    traffic = None
    tf = os.path.join(ROOT, "profiles", "flow_bwd_kernel_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")
    store_bytes = m._store_rows * ws.ld * (2 if args.precision == "bfloat16" else 4)
    return {
        "value": B / (ms * 1e-3),
        "ms_per_step": ms,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": BF16_PEAK_TFLOPS if args.precision == "bfloat16"
                     else FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / (BF16_PEAK_TFLOPS if args.precision == "bfloat16" else FP32_PEAK_TFLOPS),
                     "traffic": traffic, "kernel": "flow_bwd_kernel_h256 (adjoint of one coupling layer)",
                     "kernel_ms": ms, "flops_per_launch": fl,
                     "algorithmic_bytes_per_launch": B * (3 * d * 4 + 4) + store_bytes},
        "config": {"workload": "flow_adjoint", "batch": B, "dim": d, "bins": K, "hidden": H,
                   "precision": args.precision},
    }


WORKLOADS = {
    "nfmcmc_gmm64": (bench_nfmcmc, "min-ESS/s at R-hat<1.01 (fused NF-MCMC, 64-D 8-mode mixture)", "min-ESS/s",
                     "bf16/f32"),
    "mala_rosenbrock_128d": (bench_mala_rosenbrock, "chain-steps/s (fused MALA, 128-D Rosenbrock)", "chain-steps/s",
                             "f32"),
    "grw_rosenbrock_128d": (bench_grw_rosenbrock, "chain-steps/s (fused Gaussian random walk, 128-D Rosenbrock)",
                            "chain-steps/s", "f32"),
    "flow_train_d32": (bench_flow_train, "flow train samples/s (RQ-spline d=32, 8 layers)", "samples/s", "bf16/f32"),
    "hmc_funnels": (bench_hmc_funnels, "chain-steps/s (batched HMC, 32 funnels d in [8,256])", "chain-steps/s",
                    "f32/f64"),
    "flow_pass": (bench_flow_pass, "flow samples/s (inverse pass, d=64, 10 layers, [256,256])", "samples/s",
                  "bf16/f32"),
    "flow_adjoint": (bench_flow_adjoint, "flow adjoint samples/s (one coupling layer, d=64, [256,256], batch 10k)",
                     "samples/s", "bf16/f32"),
    "mala_gauss5d": (bench_mala_gauss5d, "chain-steps/s (fused MALA, 5-D correlated Gaussian, 1000 chains)",
                     "chain-steps/s", "f32"),
    "realnvp_train_d32": (bench_realnvp_train, "RealNVP train samples/s (d=32, 8 layers, hidden 256, batch 10k)",
                          "samples/s", "f32"),
    "pt_mala_gmm64": (bench_pt_gmm64, "tempered chain-steps/s (parallel tempering, 64-D 8-mode mixture, 5 temperatures)",
                      "tempered chain-steps/s", "f32"),
}
DEFAULT_STEPS = {"nfmcmc_gmm64": 20, "mala_rosenbrock_128d": 10048, "flow_train_d32": 50, "hmc_funnels": 1000,
                 "flow_pass": 10, "grw_rosenbrock_128d": 10048, "pt_mala_gmm64": 50,
                 "realnvp_train_d32": 50, "flow_adjoint": 50, "mala_gauss5d": 1000}
DEFAULT_WARMUP = {"nfmcmc_gmm64": 1, "mala_rosenbrock_128d": 640, "flow_train_d32": 3, "hmc_funnels": 16,
                  "flow_pass": 1, "grw_rosenbrock_128d": 640, "pt_mala_gmm64": 3,
                  "realnvp_train_d32": 3, "flow_adjoint": 3, "mala_gauss5d": 100}


def summarize_trace(path, top=25):
    """Markdown table of the kernels a rocprofv3 kernel_trace.csv records
    between the first two trace_marker() launches (the timed region)."""
    import collections
    import csv

    with open(path) as f:
        rows = sorted(csv.DictReader(f), key=lambda r: int(r["Start_Timestamp"]))
    marks = [i for i, r in enumerate(rows) if "philox_blocks_kernel" in r["Kernel_Name"]]
    if len(marks) < 2:
        raise SystemExit(f"{path}: {len(marks)} trace markers, need 2")
    a, b = marks[0], marks[1]
    win = rows[a + 1:b]
    span = (int(rows[b]["Start_Timestamp"]) - int(rows[a]["End_Timestamp"])) / 1e6
    tot = collections.defaultdict(lambda: [0, 0.0])
    busy = 0.0
    for r in win:
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
        n = r["Kernel_Name"]
        n = n.split("(")[0] if n.startswith(("void fmc", "fmc::")) else n[:100]
        tot[n][0] += 1
        tot[n][1] += dur
        busy += dur
    lines = [f"window: {len(win)} launches between the markers, {span:.1f} ms under the profiler; kernel time "
             f"{busy / 1e3:.1f} ms ({100 * busy / 1e3 / span:.1f} % of the window)", "",
             "| kernel | calls | avg (us) | total (ms) | % of kernel time |", "|---|---:|---:|---:|---:|"]
    for n, (c, t) in sorted(tot.items(), key=lambda kv: -kv[1][1])[:top]:
        lines.append(f"| `{n}` | {c} | {t / c:.1f} | {t / 1e3:.2f} | {100 * t / busy:.2f} |")
    print("\n".join(lines))


def summarize_traffic(fetch_csv, write_csv, kernel, label, out_json, n_last=0):
    """Per-launch HBM traffic of ``kernel`` from two rocprofv3 --pmc runs
    (FETCH_SIZE, then WRITE_SIZE; counter_collection.csv files), each scaled
    by what the same counter reads for the workload's 256 MiB calibration
    copy (__amd_rocclr_copyBuffer of 268,435,456 bytes each way). With
    ``n_last`` > 0 only the kernel's last n_last dispatches (the timed
    launches, after any autotuning runs) are averaged."""
    import collections
    import csv

    def per_dispatch(path, counter):
        vals = collections.defaultdict(lambda: [None, 0.0])
        with open(path) as f:
            for r in csv.DictReader(f):
                if r["Counter_Name"] != counter:
                    continue
                v = vals[int(r["Dispatch_Id"])]
                v[0] = r["Kernel_Name"]
                v[1] += float(r["Counter_Value"])
        return [vals[k] for k in sorted(vals)]

    def kib(rows, pred, last=0):
        xs = [v for n, v in rows if pred(n)]
        if not xs:
            raise SystemExit(f"no dispatch matches in the counter file")
        return xs[-last:] if last > 0 else xs

    cal = 268435456
    out = {"kernel": label}
    for key, path, counter in (("fetch", fetch_csv, "FETCH_SIZE"), ("write", write_csv, "WRITE_SIZE")):
        rows = per_dispatch(path, counter)
        k = kib(rows, lambda n: kernel in n, int(n_last))
        c = [v for v in kib(rows, lambda n: "copyBuffer" in n) if v * 1024 > cal / 4]
        out[f"{key}_size_kib_per_dispatch"] = sum(k) / len(k)
        out.setdefault("calibration", {"kernel": "__amd_rocclr_copyBuffer", "bytes_read": cal, "bytes_written": cal})
        out["calibration"][f"{key}_size_kib"] = max(c)
        out["calibration"][f"{key}_scale"] = cal / (max(c) * 1024)
    out["bytes_read_per_launch"] = out["fetch_size_kib_per_dispatch"] * 1024 * out["calibration"]["fetch_scale"]
    out["bytes_written_per_launch"] = out["write_size_kib_per_dispatch"] * 1024 * out["calibration"]["write_scale"]
    out["bytes_per_launch"] = out["bytes_read_per_launch"] + out["bytes_written_per_launch"]
    out["dispatches_averaged"] = len(k)
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


def main():
    if len(sys.argv) >= 7 and sys.argv[1] == "--traffic-summary":
        summarize_traffic(*sys.argv[2:8])
        return
    if len(sys.argv) >= 3 and sys.argv[1] == "--trace-summary":
        summarize_trace(sys.argv[2], int(sys.argv[3]) if len(sys.argv) > 3 else 25)
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="nfmcmc_gmm64", choices=list(WORKLOADS))
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--chains", type=int, default=0)
    ap.add_argument("--precision", default="bfloat16")
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--lr", type=float, default=2e-4)
    ap.add_argument("--nmax", type=int, default=100000)
    ap.add_argument("--batch", type=int, default=10000)
    ap.add_argument("--window", type=int, default=200)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--lpc", type=int, default=0)
    ap.add_argument("--nv", type=int, default=0)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--cpu-chains", type=int, default=2048)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--plain-budget", type=float, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cold", action="store_true",
                    help="nfmcmc_gmm64: skip the second run started from x0 ~ N(0, I)")
    ap.add_argument("--no-graph", action="store_true",
                    help="Rosenbrock MALA / GRW, batched HMC: launch the fused steps eagerly instead of replaying a HIP graph")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the benchmarks run on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    fn, metric, unit, dtype = WORKLOADS[args.workload]
    args.steps = args.steps if args.steps is not None else DEFAULT_STEPS[args.workload]
    args.warmup = args.warmup if args.warmup is not None else DEFAULT_WARMUP[args.workload]
    res = fn(args)
    line = {
        "metric": metric,
        "value": res["value"],
        "unit": unit,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded targets / mixtures, Philox noise)",
        "config": res["config"],
        "roofline": res["roofline"],
        "cpu_baseline": res.get("cpu_baseline"),
        "extra": res.get("extra", {}),
    }
    print(json.dumps(line, default=float))


if __name__ == "__main__":
    main()
