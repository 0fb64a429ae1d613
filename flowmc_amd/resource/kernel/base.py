"""Proposal kernel ABC — src/flowMC/resource/kernel/base.py:9-27.

Besides the reference's per-chain ``kernel`` signature, every proposal offers
``sample_steps``: the batched, buffer-writing entry point the step strategies
call (the reference vmaps ``kernel`` over chains and scans it over steps,
src/flowMC/strategy/take_steps.py:127-180).
"""

from __future__ import annotations

import ctypes as C
from abc import abstractmethod
from dataclasses import dataclass
from typing import Optional

import torch

from ... import _native
from ... import random as frandom
from ..base import Resource

# fmc_kernel_kind of the C ABI
KERNEL_MALA, KERNEL_HMC, KERNEL_GRW = 0, 1, 2


@dataclass
class StepOutputs:
    """Device buffers a strategy writes (rows [C, T, ...] along dim 1) and the
    row at which the first thinned step lands (already clamped like
    jax.lax.dynamic_update_slice)."""

    positions: Optional[torch.Tensor]
    log_probs: Optional[torch.Tensor]
    accepts: Optional[torch.Tensor]
    cursor: int = 0

    @property
    def buf_len(self) -> int:
        for t in (self.positions, self.log_probs, self.accepts):
            if t is not None:
                return int(t.shape[1])
        return 1


def rows_written(n_steps: int, thinning: int) -> int:
    """Rows positions[:, ::thinning] has (take_steps.py:134-138)."""
    return (n_steps + thinning - 1) // thinning


def steps_needed(n_steps: int, thinning: int) -> int:
    """Only steps up to the last thinned row influence buffers and the returned
    position (take_steps.py:144 returns positions[:, ::thinning][:, -1])."""
    return (rows_written(n_steps, thinning) - 1) * thinning + 1 if n_steps > 0 else 0


class ProposalBase(Resource):
    @abstractmethod
    def __init__(self):
        """Initialize the sampler class."""

    @abstractmethod
    def kernel(self, rng_key, position, log_prob, logpdf, data):
        """One proposal step for one chain ([d]) or a batch of chains ([C, d])."""

    @abstractmethod
    def sample_steps(self, rng_key, initial_position: torch.Tensor, logpdf, data, n_steps: int,
                     thinning: int, outputs: StepOutputs, noise=None, log_u=None) -> torch.Tensor:
        """Run ``n_steps`` for all chains, write the thinned rows into ``outputs``
        and return the position of the last thinned row."""


    # (fmc_kernel_kind, step_size, n_leapfrog, metric [d] or None[, momentum
    # factor [d, d] or None]) of the fused tempered launch, or None when the
    # proposal has no fused local step
    def _tempered_launch(self, n_dim: int, device):
        return None

    def sample_tempered(self, rng_key, positions: torch.Tensor, tempered_logpdf, temperatures: torch.Tensor, data,
                        n_steps: int, noise=None, log_u=None, accepts: Optional[torch.Tensor] = None):
        """``n_steps`` local steps of every (chain, temperature) row —
        ParallelTempering._ensemble_step / _individal_step
        (src/flowMC/strategy/parallel_tempering.py:135-290), batched over chains.

        positions [C, T, d]; row c*T + t targets
        ``tempered_logpdf.tempered_log_pdf(temperatures[t], x)``. Returns the new
        positions [C, T, d] and their tempered log-densities [C, T]; ``accepts``
        [C*T, n_steps] (optional) receives the step decisions. Noise, when
        injected, is laid out over the rows: noise [n_steps, C*T, d],
        log_u [n_steps, C*T].
        """
        from ..logPDF import TemperedPDF, TemperedRows

        _native.require_gpu()
        x = torch.as_tensor(positions, dtype=torch.float32)
        x = (x if x.is_cuda else x.cuda()).contiguous().clone()
        C_, T, d = x.shape
        temps = torch.as_tensor(temperatures, dtype=torch.float32)
        temps = (temps if temps.is_cuda else temps.to(x.device)).contiguous()
        if temps.numel() != T:
            raise ValueError(f"{temps.numel()} temperatures for {T} tempered slots")
        rows = x.reshape(C_ * T, d)
        if noise is not None:
            noise = noise.to(device=x.device, dtype=torch.float32).contiguous()
            log_u = log_u.to(device=x.device, dtype=torch.float32).contiguous()
            if noise.shape[0] < n_steps or log_u.shape[0] < n_steps:
                raise ValueError("injected noise must cover every step that is run")
        launch = None
        if isinstance(tempered_logpdf, TemperedPDF) and tempered_logpdf.fused_tempering:
            launch = self._tempered_launch(d, x.device)
        if launch is None:
            out = StepOutputs(None, None, accepts, 0)
            rows = self.sample_steps(rng_key, rows, TemperedRows(tempered_logpdf, temps), data, n_steps, 1, out,
                                     noise=noise, log_u=log_u)
            lp = self.last_log_prob
            return rows.reshape(C_, T, d), lp.reshape(C_, T)
        kind, step_size, n_leapfrog, metric, *rest = launch
        chol = rest[0] if rest else None
        desc = tempered_logpdf.log_pdf.descriptor()
        lp = torch.empty(C_ * T, device=x.device, dtype=torch.float32)
        k0, k1 = frandom.key_words(rng_key)
        cfg = getattr(self, "launch_cfg", None)
        cfg = C.byref(cfg) if cfg is not None else None
        st = _native.stream_ptr()
        K = max(1, int(getattr(self, "fused_steps", 64)))
        acc_len = int(accepts.shape[1]) if accepts is not None else 0
        for s0 in range(0, n_steps, K):
            ns = min(K, n_steps - s0)
            nz = noise[s0:s0 + ns] if noise is not None else None
            lu = log_u[s0:s0 + ns] if log_u is not None else None
            _native.call("fmc_tempered_run", st, kind, C.byref(desc), _native.dptr(rows, name="positions"),
                         _native.dptr(lp, name="logp"), C_, T, _native.dptr(temps, name="temperatures", numel=T), ns,
                         s0, float(step_size), int(n_leapfrog),
                         _native.dptr(metric, name="metric", numel=d) if metric is not None else None,
                         _native.dptr(chol, name="momentum factor", numel=d * d) if chol is not None else None, k0, k1,
                         _native.dptr(nz, name="noise") if nz is not None else None,
                         _native.dptr(lu, name="log_u") if lu is not None else None,
                         _native.dptr(accepts, name="acceptance buffer") if accepts is not None else None, acc_len,
                         cfg)
        self.last_log_prob = lp
        return rows.reshape(C_, T, d), lp.reshape(C_, T)
