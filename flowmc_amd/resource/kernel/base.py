"""Proposal kernel ABC — src/flowMC/resource/kernel/base.py:9-27.

Besides the reference's per-chain ``kernel`` signature, every proposal offers
``sample_steps``: the batched, buffer-writing entry point the step strategies
call (the reference vmaps ``kernel`` over chains and scans it over steps,
src/flowMC/strategy/take_steps.py:127-180).
"""

from __future__ import annotations

from abc import abstractmethod
from dataclasses import dataclass
from typing import Optional

import torch

from ..base import Resource


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
