"""Benchmark of the flowMC hot path on one MI355X.

Default workload (chain-steps/s): fused MALA on the 128-D Rosenbrock target
log p = -sum_i [100 (x_{i+1} - x_i^2)^2 + (1 - x_i)^2] / 20, 16384 chains,
step 0.01, 64 steps fused per launch, thinning 1, every step's position /
log-prob / acceptance row written to device buffers of the full run length
(as TakeSerialSteps does). One timed "step" = one MALA step of all chains.

Prints ONE JSON line (see the keys below). ``--gpus N`` with N > 1 exits: the
build targets a single GPU.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
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


def bench_mala_rosenbrock(args):
    import torch

    import flowmc_amd._native as N
    from flowmc_amd import random as fr
    from flowmc_amd import targets as TG

    C_, d, dt, K = args.chains, args.dim, args.step_size, args.fused
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
    sp = stream.cuda_stream
    cfg = N.fmc_launch_cfg(args.lpc, args.nv, args.block, 0) if args.lpc else None
    cfgp = C.byref(cfg) if cfg is not None else None

    def run(n_steps, key):
        k0, k1 = fr.key_words(key)
        for s0 in range(0, n_steps, K):
            ns = min(K, n_steps - s0)
            N.call("fmc_mala_run", sp, C.byref(desc), x.data_ptr(), lp.data_ptr(), C_, ns, s0, dt, k0, k1, None, None,
                   pos.data_ptr(), lps.data_ptr(), acc.data_ptr(), T, 0, 1, cfgp)

    run(args.warmup, fr.PRNGKey(1))
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    run(steps, fr.PRNGKey(2))
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    n_launch = (steps + K - 1) // K
    # algorithmic HBM bytes: per chain-step the position row (4d), log-prob (4)
    # and acceptance (4); per launch and chain the state read + write (8d + 8)
    bytes_total = C_ * (steps * (4 * d + 8) + n_launch * (8 * d + 8))
    achieved = bytes_total / (gpu_ms * 1e-3) / 1e9
    accept = float(acc[:, :steps].mean())
    # FLOP-side count per chain-step: Rosenbrock value+grad ~17d, MALA update ~12d
    flops = C_ * steps * (17 * d + 12 * d)
    return {
        "value": C_ * steps / wall,
        "ms_per_step": wall * 1e3 / steps,
        "kernel_ms_per_launch": gpu_ms / n_launch,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "fp32_tflops": flops / (gpu_ms * 1e-3) / 1e12,
        "acceptance": accept,
        "bytes_per_launch": bytes_total / n_launch,
    }


def cpu_baseline_mala_rosenbrock(args):
    """The FP32 CPU restatement (oracle, 2 value_and_grad per step as MALA.py:59,71-73)
    on a bounded sample of chains and steps, timed on the host cores."""
    import torch

    from oracle import chains as OC
    from oracle import targets as OT

    torch.set_num_threads(os.cpu_count() or 1)
    C_, d, dt = args.cpu_chains, args.dim, args.step_size
    g = torch.Generator().manual_seed(0)
    x = torch.randn(C_, d, generator=g)
    fn = lambda z: OT.rosenbrock(z)
    steps = 0
    OC.mala_step(x, fn, dt, torch.randn(C_, d, generator=g), torch.log(torch.rand(C_, generator=g)))
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < args.cpu_seconds:
        xi = torch.randn(C_, d, generator=g)
        lu = torch.log(torch.rand(C_, generator=g))
        x, _, _, _ = OC.mala_step(x, fn, dt, xi, lu)
        steps += 1
    el = time.perf_counter() - t0
    return {"value": C_ * steps / el, "unit": "chain-steps/s", "cores": torch.get_num_threads(), "kind": "port",
            "sample": f"{C_} chains x {steps} MALA steps, d={d}, FP32 torch-CPU oracle restatement "
                      f"({cpu_model()})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10048)
    ap.add_argument("--warmup", type=int, default=640)
    ap.add_argument("--chains", type=int, default=16384)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--step-size", type=float, default=0.01)
    ap.add_argument("--fused", type=int, default=64)
    ap.add_argument("--lpc", type=int, default=0)
    ap.add_argument("--nv", type=int, default=0)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--cpu-chains", type=int, default=16384)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    res = bench_mala_rosenbrock(args)
    cpu = None if args.no_cpu_baseline else cpu_baseline_mala_rosenbrock(args)
    line = {
        "metric": "chain-steps/s (fused MALA, 128-D Rosenbrock)",
        "value": res["value"],
        "unit": "chain-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (x0 ~ N(0, I), in-kernel Philox noise)",
        "config": {"workload": "mala_rosenbrock_128d", "chains": args.chains, "dim": args.dim,
                   "step_size": args.step_size, "fused_steps_per_launch": args.fused, "thinning": 1},
        "roofline": res["roofline"],
        "cpu_baseline": cpu,
        "extra": {"kernel_ms_per_launch": res["kernel_ms_per_launch"], "fp32_tflops": res["fp32_tflops"],
                  "acceptance": res["acceptance"], "bytes_per_launch": res["bytes_per_launch"]},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
