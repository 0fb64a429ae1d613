"""Optimizer resource — src/flowMC/resource/optimizer.py:6-35.

``optim`` is a gradient transformation description executed by the fused
device kernel fmc_adamw_clip (global-norm clip + AdamW on the flat master
parameter vector); ``optim_state`` holds the moments and the step count.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import torch

from .base import Resource


@dataclass
class AdamWClip:
    """optax.chain(clip_by_global_norm(max_norm), adamw(lr, b1, b2, eps, weight_decay)).
    max_norm None = no clipping; weight_decay 0 = plain Adam (optax.adam)."""

    learning_rate: float = 1e-3
    b1: float = 0.9
    b2: float = 0.999
    eps: float = 1e-8
    weight_decay: float = 1e-4
    max_norm: Optional[float] = 1.0

    def init(self, model) -> "AdamState":
        n = model.theta.numel()
        dev = model.theta.device
        return AdamState(torch.zeros(n, dtype=torch.float32, device=dev),
                         torch.zeros(n, dtype=torch.float32, device=dev), 0)


@dataclass
class AdamState:
    mu: torch.Tensor
    nu: torch.Tensor
    count: int = 0

    def clone(self) -> "AdamState":
        return AdamState(self.mu.clone(), self.nu.clone(), self.count)


def adamw(learning_rate=1e-3, b1=0.9, b2=0.999, eps=1e-8, weight_decay=1e-4, max_norm=None) -> AdamWClip:
    return AdamWClip(learning_rate, b1, b2, eps, weight_decay, max_norm)


def adam(learning_rate=1e-3, b1=0.9, b2=0.999, eps=1e-8) -> AdamWClip:
    """optax.adam (as used by test/integration/test_normalizingFlow.py:53)."""
    return AdamWClip(learning_rate, b1, b2, eps, 0.0, None)


class Optimizer(Resource):
    optim: AdamWClip
    optim_state: AdamState

    def __repr__(self):
        return "Optimizer"

    def __init__(self, model, learning_rate: float = 1e-3, momentum: float = 0.9):
        # optimizer.py:19-22: clip_by_global_norm(1.0) then adamw(lr, b1=momentum)
        self.optim = AdamWClip(learning_rate=learning_rate, b1=momentum, weight_decay=1e-4, max_norm=1.0)
        self.optim_state = self.optim.init(model)

    def __call__(self, params, grads):
        raise NotImplementedError

    def print_parameters(self):
        raise NotImplementedError

    def save_resource(self, path: str):
        raise NotImplementedError

    def load_resource(self, path: str):
        raise NotImplementedError
