"""optax.chain(clip_by_global_norm(max_norm), adamw(lr, b1, b2, eps, weight_decay))
on a flat parameter vector (src/flowMC/resource/optimizer.py:19-22), FP64.

optax 0.2.4 semantics restated: the global norm is taken over all gradient
leaves; if norm >= max_norm every update is g / norm * max_norm; Adam moments
with bias correction by 1 - b^t (t counted from 1); the decoupled weight
decay wd * param is added before scaling by -lr; every inexact leaf is
decayed (parity unpinned for the stop-gradient leaves: optax is absent).
"""

from __future__ import annotations

import torch


def adamw_clip_step(theta, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8, wd=1e-4, max_norm=1.0):
    g = grad
    if max_norm is not None:
        norm = torch.sqrt((g * g).sum())
        if not bool(norm < max_norm):
            g = g / norm * max_norm
    m = (1 - b1) * g + b1 * m
    v = (1 - b2) * g * g + b2 * v
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    upd = mh / (torch.sqrt(vh) + eps) + wd * theta
    return theta - lr * upd, m, v
