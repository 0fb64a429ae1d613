"""Philox4x32-10 and the uniform / normal conversions used by the chain kernels.

Integer part is bit-exact (numpy uint32/uint64); the Box–Muller transform is
evaluated in FP64 from the same FP32 uniforms the kernels use.
Published algorithm: Salmon et al., SC'11 (Random123).
"""

from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)

# stream ids (4th counter word)
RNG_MALA_NOISE = 0
RNG_MALA_UNIFORM = 1
RNG_HMC_MOMENTUM = 2
RNG_HMC_UNIFORM = 3
RNG_NF_UNIFORM = 4
RNG_FLOW_BASE = 5
RNG_DIRECT_NORMAL = 6
RNG_GRW_NOISE = 7
RNG_GRW_UNIFORM = 8
RNG_PT_SWAP = 9


def philox4x32_10(ctr, key) -> np.ndarray:
    """ctr: array [..., 4] uint32; key: (k0, k1). Returns [..., 4] uint32."""
    c = np.asarray(ctr, dtype=np.uint32)
    c0, c1, c2, c3 = (c[..., i].astype(np.uint64) for i in range(4))
    k0 = np.uint32(key[0])
    k1 = np.uint32(key[1])
    with np.errstate(over="ignore"):
        for _ in range(10):
            p0 = M0 * c0
            p1 = M1 * c2
            hi0, lo0 = p0 >> np.uint64(32), p0 & MASK32
            hi1, lo1 = p1 >> np.uint64(32), p1 & MASK32
            n0 = hi1 ^ c1 ^ np.uint64(k0)
            n2 = hi0 ^ c3 ^ np.uint64(k1)
            c0, c1, c2, c3 = n0, lo1, n2, lo0
            k0 = np.uint32(np.uint64(k0) + np.uint64(W0) & MASK32)
            k1 = np.uint32(np.uint64(k1) + np.uint64(W1) & MASK32)
    return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


def u01(bits: np.ndarray) -> np.ndarray:
    """Open-interval uniform exactly as the kernels form it (exact in FP32)."""
    b = np.asarray(bits, dtype=np.uint32)
    return ((b >> np.uint32(9)).astype(np.float64) + 0.5) * 2.0 ** -23


def normal4(key, chunk, chain, step, sid) -> np.ndarray:
    """Four normals for dims 4*chunk..4*chunk+3 (Box–Muller on word pairs)."""
    chunk, chain, step = np.broadcast_arrays(np.asarray(chunk), np.asarray(chain), np.asarray(step))
    ctr = np.stack([chunk, chain, step, np.full_like(chunk, sid)], axis=-1).astype(np.uint32)
    b = philox4x32_10(ctr, key)
    r0 = np.sqrt(-2.0 * np.log(u01(b[..., 0])))
    r1 = np.sqrt(-2.0 * np.log(u01(b[..., 2])))
    a0 = 2.0 * np.pi * u01(b[..., 1])
    a1 = 2.0 * np.pi * u01(b[..., 3])
    return np.stack([r0 * np.cos(a0), r0 * np.sin(a0), r1 * np.cos(a1), r1 * np.sin(a1)], axis=-1)


def normals(key, n_chains: int, n_dim: int, n_steps: int, step0: int, sid: int) -> np.ndarray:
    """[n_steps, n_chains, n_dim] normals of (chain, step0+s, dim)."""
    chunks = (n_dim + 3) // 4
    s, c, k = np.meshgrid(np.arange(n_steps) + step0, np.arange(n_chains), np.arange(chunks), indexing="ij")
    z = normal4(key, k, c, s, sid)  # [S, C, chunks, 4]
    return z.reshape(n_steps, n_chains, chunks * 4)[..., :n_dim]


def log_uniforms(key, n_chains: int, n_steps: int, step0: int, sid: int) -> np.ndarray:
    s, c = np.meshgrid(np.arange(n_steps) + step0, np.arange(n_chains), indexing="ij")
    ctr = np.stack([np.zeros_like(c), c, s, np.full_like(c, sid)], axis=-1).astype(np.uint32)
    b = philox4x32_10(ctr, key)
    # kernels take logf of the FP32 uniform; the uniform itself is exact
    return np.log(u01(b[..., 0]))
