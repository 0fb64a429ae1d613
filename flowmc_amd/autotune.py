"""Small launch-shape autotuner.

Times each candidate configuration of a kernel launch with HIP events on the
current stream (after one untimed warm-up) and keeps the fastest; results
are cached per problem signature for the life of the process.
"""

from __future__ import annotations

from typing import Callable, Hashable, Iterable

import torch

_cache: dict[Hashable, object] = {}


def best_config(signature: Hashable, candidates: Iterable, run: Callable[[object], None], reps: int = 3):
    """Fastest candidate for ``signature``; ``run(cfg)`` launches the work once."""
    if signature in _cache:
        return _cache[signature]
    best, best_t = None, float("inf")
    stream = torch.cuda.current_stream()
    for cfg in candidates:
        try:
            run(cfg)  # warm-up (and validation: invalid layouts raise)
        except Exception:
            continue
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record(stream)
        for _ in range(reps):
            run(cfg)
        ev1.record(stream)
        ev1.synchronize()
        t = ev0.elapsed_time(ev1) / reps
        if t < best_t:
            best, best_t = cfg, t
    if best is None:
        raise RuntimeError(f"no valid launch configuration for {signature}")
    _cache[signature] = best
    return best


def clear():
    _cache.clear()


def chain_layouts(d: int) -> list[tuple[int, int]]:
    """(lanes_per_chain, vec_per_lane) pairs that hold d dims with < 2x padding."""
    out = []
    for nv in (1, 2, 4, 8):
        lpc = 1
        while lpc * 4 * nv < d and lpc < 64:
            lpc *= 2
        if lpc * 4 * nv >= d:
            out.append((lpc, nv))
    return out
