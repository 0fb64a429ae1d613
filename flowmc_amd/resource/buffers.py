"""Buffer resource — src/flowMC/resource/buffers.py:11-64.

Device-resident storage initialised to -inf. ``update_buffer`` follows
jax.lax.dynamic_update_slice_in_dim (buffers.py:39-41): the start index is
clamped to [0, len - n] along ``cursor_dim`` so the update always fits. The
write is done in place (the reference rebuilds the whole array per update).
"""

from __future__ import annotations

import numpy as np
import torch

from .base import Resource


def clamp_start(start: int, n_update: int, n_total: int) -> int:
    """Start index jax.lax.dynamic_update_slice actually uses."""
    if n_update > n_total:
        raise ValueError(f"update of length {n_update} does not fit a buffer of length {n_total}")
    return max(0, min(int(start), n_total - n_update))


def default_device():
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")


class Buffer(Resource):
    name: str
    data: torch.Tensor
    cursor: int = 0
    cursor_dim: int = 0

    def __repr__(self):
        return "Buffer " + self.name + " with shape " + str(tuple(self.data.shape))

    @property
    def shape(self):
        return tuple(self.data.shape)

    def __init__(self, name: str, shape: tuple, cursor_dim: int = 0, device=None,
                 dtype: torch.dtype = torch.float32):
        self.cursor_dim = cursor_dim
        self.cursor = 0
        self.name = name
        dev = default_device() if device is None else torch.device(device)
        self.data = torch.full(tuple(shape), -float("inf"), dtype=dtype, device=dev)

    def __call__(self):
        return self.data

    def update_buffer(self, updates, start: int = 0):
        """Write ``updates`` at ``start`` along ``cursor_dim`` (clamped like JAX)."""
        u = torch.as_tensor(updates, dtype=self.data.dtype, device=self.data.device)
        ax = self.cursor_dim
        if u.ndim != self.data.ndim:
            raise ValueError(f"update rank {u.ndim} != buffer rank {self.data.ndim}")
        s = clamp_start(start, u.shape[ax], self.data.shape[ax])
        idx = [slice(None)] * self.data.ndim
        idx[ax] = slice(s, s + u.shape[ax])
        for a in range(self.data.ndim):
            if a != ax and u.shape[a] != self.data.shape[a]:
                raise ValueError(f"update shape {tuple(u.shape)} incompatible with buffer {self.shape}")
        self.data[tuple(idx)] = u

    def print_parameters(self):
        print(
            f"Buffer: {self.name} with shape {self.shape} and cursor"
            f" {self.cursor} at dimension {self.cursor_dim}"
        )

    def get_distribution(self, n_bins: int = 100):
        return np.histogram(self.data.detach().float().cpu().numpy().flatten(), bins=n_bins)

    def save_resource(self, path: str):
        """buffers.py:52-57: ``np.savez(path + name, name=name, data=data)`` (an
        ``.npz`` with a 0-d string array and the data array; no pickle)."""
        np.savez(path + self.name, name=self.name, data=self.data.detach().cpu().numpy())

    def load_resource(self, path: str) -> "Buffer":
        """buffers.py:59-64: a new Buffer named and shaped by the file, with
        the default cursor (0) and cursor_dim (0), as the reference's
        ``Buffer(data["name"], buffer.shape)``. Pickled files are refused."""
        f = np.load(path, allow_pickle=False)
        arr = f["data"]
        result = Buffer(str(f["name"]), arr.shape, device=self.data.device)
        result.data = torch.as_tensor(arr, dtype=self.data.dtype, device=self.data.device)
        return result
