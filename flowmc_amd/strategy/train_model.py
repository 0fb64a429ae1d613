"""TrainModel — src/flowMC/strategy/train_model.py:11-112."""

from __future__ import annotations

import torch

from .. import _native
from .. import random as frandom
from ..resource.buffers import Buffer
from ..resource.model.nf_model.base import NFModel
from ..resource.optimizer import Optimizer
from .base import Strategy


def select_training_data(data: torch.Tensor, history_window: int, n_max_examples: int, key) -> torch.Tensor:
    """train_model.py:66-80: keep rows whose every entry is finite, reshape them
    to (n_chains, -1, n_dims), take ``[:, -history_window:]``, then draw
    ``n_max_examples`` rows with replacement.

    Only the drawn rows are gathered: the finite-row mask comes from one pass
    of fmc_finite_rows over the buffer, its running count locates the k-th
    finite row (chain-major order, as the reference's boolean indexing lists
    them), and the slice is applied to ranks. Same rows, same order, same
    errors as filtering and reshaping the whole buffer, without copying it.
    """
    n_chains, n_dims = data.shape[0], data.shape[-1]
    flat = data.reshape(-1, n_dims)
    if flat.dtype != torch.float32 or not flat.is_contiguous():
        flat = flat.to(torch.float32).contiguous()
    mask = torch.empty(flat.shape[0], dtype=torch.int32, device=flat.device)
    _native.call("fmc_finite_rows", _native.stream_ptr(), _native.dptr(flat, name="training buffer"),
                 flat.shape[0], n_dims, _native.dptr(mask, torch.int32))
    rank = torch.cumsum(mask, 0, dtype=torch.int64)
    n_finite = int(rank[-1]) if rank.numel() else 0
    if n_finite % n_chains != 0:
        raise ValueError("chains hold different numbers of finite rows; cannot reshape to (n_chains, -1, n_dims)")
    m = n_finite // n_chains
    start, stop, _ = slice(-history_window, None).indices(m)
    w = max(0, stop - start)
    if n_chains * w == 0:
        raise ValueError("no finite rows in the training buffer")
    idx = frandom.choice(key, n_chains * w, n_max_examples, device=flat.device)
    k = (idx // w) * m + start + idx % w  # rank of the drawn row among the finite rows
    return flat[torch.searchsorted(rank, k + 1)]


class TrainModel(Strategy):
    def __repr__(self):
        return "Train " + self.model_resource

    def __init__(self, model_resource: str, data_resource: str, optimizer_resource: str, loss_buffer_name: str = "",
                 n_epochs: int = 100, batch_size: int = 64, n_max_examples: int = 10000, history_window: int = 100,
                 verbose: bool = False):
        self.model_resource = model_resource
        self.data_resource = data_resource
        self.optimizer_resource = optimizer_resource
        self.loss_buffer_name = loss_buffer_name
        self.n_epochs = n_epochs
        self.batch_size = batch_size
        self.n_max_examples = n_max_examples
        self.verbose = verbose
        self.history_window = history_window

    def __call__(self, rng_key, resources, initial_position, data):
        model = resources[self.model_resource]
        assert isinstance(model, NFModel), "Target resource must be a NFModel"
        data_resource = resources[self.data_resource]
        assert isinstance(data_resource, Buffer), "Data resource must be a buffer"
        optimizer = resources[self.optimizer_resource]
        assert isinstance(optimizer, Optimizer), "Optimizer resource must be an optimizer"
        rng_key, subkey = frandom.split(rng_key)
        training_data = select_training_data(data_resource.data, self.history_window, self.n_max_examples, subkey)
        rng_key, subkey = frandom.split(rng_key)
        if self.verbose:
            print("Training model")
            print(f"Training data shape: {tuple(training_data.shape)}")
            print(f"n_epochs: {self.n_epochs}")
            print(f"batch_size: {self.batch_size}")
        rng_key, model, optim_state, loss_values = model.train(rng=subkey, data=training_data, optim=optimizer.optim,
                                                               state=optimizer.optim_state, num_epochs=self.n_epochs,
                                                               batch_size=self.batch_size, verbose=self.verbose)
        if self.loss_buffer_name != "":
            loss_buffer = resources[self.loss_buffer_name]
            assert isinstance(loss_buffer, Buffer), "Loss buffer resource must be a buffer"
            loss_buffer.update_buffer(loss_values, start=loss_buffer.cursor)
            loss_buffer.cursor += len(loss_values)
            resources[self.loss_buffer_name] = loss_buffer
        optimizer.optim_state = optim_state
        resources[self.model_resource] = model
        resources[self.optimizer_resource] = optimizer
        return rng_key, resources, initial_position
