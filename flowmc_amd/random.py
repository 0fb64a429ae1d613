"""Explicit-key random numbers (the role of jax.random in the reference).

A key is a numpy ``uint32[2]``. Keys are split on the host with one
Philox4x32-10 block per child; the device kernels consume the key directly
and derive per-chain / per-step streams from the counter
(dim_chunk, chain, step, stream_id) — see include/flowmc.h.

This replaces the jax.random key plumbing of the reference
(src/flowMC/strategy/take_steps.py:71-72, MALA.py:66, HMC.py:128,
NF_proposal.py:41,99). Bit-level equality with JAX's threefry streams is not
attempted; determinism is "same key -> same draws".
"""

from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from . import _native

_M0, _M1 = 0xD2511F53, 0xCD9E8D57
_W0, _W1 = 0x9E3779B9, 0xBB67AE85
_MASK = 0xFFFFFFFF

# counter word 3 used for host-side key derivation and direct draws
_SID_SPLIT = 0x53504C54  # "SPLT"
SID_DIRECT_NORMAL = 6


def _philox(ctr: Sequence[int], key: Sequence[int]) -> tuple[int, int, int, int]:
    c0, c1, c2, c3 = (int(v) & _MASK for v in ctr)
    k0, k1 = int(key[0]) & _MASK, int(key[1]) & _MASK
    for _ in range(10):
        p0 = _M0 * c0
        p1 = _M1 * c2
        c0, c1, c2, c3 = ((p1 >> 32) ^ c1 ^ k0) & _MASK, p1 & _MASK, ((p0 >> 32) ^ c3 ^ k1) & _MASK, p0 & _MASK
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return c0, c1, c2, c3


def PRNGKey(seed: int) -> np.ndarray:
    """Key from an integer seed (same (hi, lo) word layout as jax.random.PRNGKey)."""
    seed = int(seed)
    return np.array([(seed >> 32) & _MASK, seed & _MASK], dtype=np.uint32)


def _as_key(key) -> np.ndarray:
    k = np.asarray(key, dtype=np.uint32).reshape(-1)
    if k.shape != (2,):
        raise ValueError(f"a key is uint32[2], got shape {np.asarray(key).shape}")
    return k


def split(key, num: int = 2) -> np.ndarray:
    """num child keys, shape [num, 2]."""
    k = _as_key(key)
    out = np.empty((num, 2), dtype=np.uint32)
    for i in range(num):
        b = _philox((i, 0, 0, _SID_SPLIT), k)
        out[i, 0], out[i, 1] = b[0], b[1]
    return out


def fold_in(key, data: int) -> np.ndarray:
    k = _as_key(key)
    b = _philox((int(data) & _MASK, (int(data) >> 32) & _MASK, 1, _SID_SPLIT), k)
    return np.array([b[0], b[1]], dtype=np.uint32)


def key_words(key) -> tuple[int, int]:
    k = _as_key(key)
    return int(k[0]), int(k[1])


def normal(key, shape, device="cuda", dtype=torch.float32) -> torch.Tensor:
    """Standard normals of the given shape, drawn on the device by fmc_rng_normal."""
    if isinstance(shape, int):
        shape = (shape,)
    shape = tuple(int(s) for s in shape)
    n = int(np.prod(shape)) if shape else 1
    d = shape[-1] if shape else 1
    rows = n // d if d else 0
    out = torch.empty(shape, device=device, dtype=torch.float32)
    if n:
        _native.require_gpu()
        k0, k1 = key_words(key)
        _native.call("fmc_rng_normal", _native.stream_ptr(), k0, k1, rows, d, 1, 0, SID_DIRECT_NORMAL,
                     _native.dptr(out, name="out"))
    return out.to(dtype)


def torch_generator(key, device="cuda") -> torch.Generator:
    """A torch.Generator seeded from a key (index sampling: choice / permutation)."""
    k0, k1 = key_words(key)
    g = torch.Generator(device=device)
    g.manual_seed((k0 << 32) | k1)
    return g


def permutation(key, n: int, device="cuda") -> torch.Tensor:
    return torch.randperm(n, generator=torch_generator(key, device), device=device)


def choice(key, n: int, size: int, device="cuda") -> torch.Tensor:
    """Indices drawn uniformly with replacement (jax.random.choice(..., replace=True))."""
    return torch.randint(0, n, (size,), generator=torch_generator(key, device), device=device)
