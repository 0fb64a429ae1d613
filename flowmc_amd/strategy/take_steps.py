"""TakeSteps / TakeSerialSteps / TakeGroupSteps — src/flowMC/strategy/take_steps.py:14-206.

The reference vmaps ``kernel.kernel`` over chains, scans it over steps, thins
the stacked outputs and copies them into the buffers. Here the proposal's
``sample_steps`` runs the whole batch on the device and writes the thinned
rows straight into the buffers at the (clamped) cursor.
"""

from __future__ import annotations

from abc import abstractmethod

import torch

from .. import random as frandom
from ..resource.buffers import Buffer, clamp_start
from ..resource.kernel.base import ProposalBase, StepOutputs, rows_written
from ..resource.states import State
from .base import Strategy


class TakeSteps(Strategy):
    logpdf_name: str
    kernel_name: str
    state_name: str
    buffer_names: list
    n_steps: int
    current_position: int
    thinning: int
    chain_batch_size: int
    verbose: bool

    def __init__(self, logpdf_name: str, kernel_name: str, state_name: str, buffer_names: list, n_steps: int,
                 thinning: int = 1, chain_batch_size: int = 0, verbose: bool = False):
        self.logpdf_name = logpdf_name
        self.kernel_name = kernel_name
        self.state_name = state_name
        self.buffer_names = buffer_names
        self.n_steps = n_steps
        self.current_position = 0
        self.thinning = thinning
        # The device kernels keep no per-step intermediates in memory, so
        # batching chains (take_steps.py:105-125) would not change results or
        # memory; the value is kept for API compatibility.
        self.chain_batch_size = chain_batch_size
        self.verbose = verbose

    def set_current_position(self, current_position: int):
        self.current_position = current_position

    def _resolve(self, resources):
        assert isinstance(state_resource := resources[self.state_name], State), "State resource must be a State"
        assert isinstance(position_buffer_name := state_resource.data[self.buffer_names[0]], str), (
            "Position buffer resource name must be a string")
        assert isinstance(log_prob_buffer_name := state_resource.data[self.buffer_names[1]], str), (
            "Log probability buffer resource name must be a string")
        assert isinstance(acceptance_buffer_name := state_resource.data[self.buffer_names[2]], str), (
            "Acceptance buffer resource name must be a string")
        assert isinstance(position_buffer := resources[position_buffer_name], Buffer), (
            "Position buffer resource must be a Buffer")
        assert isinstance(log_prob_buffer := resources[log_prob_buffer_name], Buffer), (
            "Log probability buffer resource must be a Buffer")
        assert isinstance(acceptance_buffer := resources[acceptance_buffer_name], Buffer), (
            "Acceptance buffer resource must be a Buffer")
        return position_buffer, log_prob_buffer, acceptance_buffer

    def outputs(self, resources, n_chains: int, n_dim: int) -> StepOutputs:
        pb, lb, ab = self._resolve(resources)
        for b, rank in ((pb, 3), (lb, 2), (ab, 2)):
            if b.cursor_dim != 1 or b.data.ndim != rank or b.data.shape[0] != n_chains:
                raise ValueError(f"{b!r}: step buffers must be [n_chains, n_steps(, n_dim)] with cursor_dim=1")
            if not (b.data.is_cuda and b.data.dtype == torch.float32 and b.data.is_contiguous()):
                raise ValueError(f"{b!r}: step buffers must be contiguous float32 device tensors")
        if pb.data.shape[2] != n_dim:
            raise ValueError(f"position buffer last dim {pb.data.shape[2]} != n_dim {n_dim}")
        n_rows = rows_written(self.n_steps, self.thinning)
        T = pb.data.shape[1]
        if lb.data.shape[1] != T or ab.data.shape[1] != T:
            raise ValueError("position / log_prob / acceptance buffers have different lengths")
        start = clamp_start(self.current_position, n_rows, T)
        return StepOutputs(pb.data, lb.data, ab.data, start)

    @abstractmethod
    def run(self, kernel, subkey, initial_position, logpdf, data, outputs: StepOutputs):
        raise NotImplementedError

    def __call__(self, rng_key, resources, initial_position, data):
        rng_key, subkey = frandom.split(rng_key)  # take_steps.py:71
        kernel = resources[self.kernel_name]
        logpdf = resources[self.logpdf_name]
        assert isinstance(kernel, ProposalBase), "Kernel resource must be a ProposalBase"
        x0 = torch.as_tensor(initial_position, dtype=torch.float32)
        if x0.ndim == 1:
            x0 = x0[None]
        if not x0.is_cuda:
            x0 = x0.cuda()
        n_chains, n_dim = x0.shape
        out = self.outputs(resources, n_chains, n_dim)
        last = self.run(kernel, subkey, x0, logpdf, data, out)
        self.current_position += self.n_steps // self.thinning  # take_steps.py:143
        return rng_key, resources, last


class TakeSerialSteps(TakeSteps):
    """Sequential steps of a local kernel (take_steps.py:147-180)."""

    def run(self, kernel, subkey, initial_position, logpdf, data, outputs):
        return kernel.sample_steps(subkey, initial_position, logpdf, data, self.n_steps, self.thinning, outputs)


class TakeGroupSteps(TakeSteps):
    """All steps drawn at once, e.g. normalizing-flow proposals (take_steps.py:183-206)."""

    def run(self, kernel, subkey, initial_position, logpdf, data, outputs):
        return kernel.sample_steps(subkey, initial_position, logpdf, {**(data or {}), "n_steps": self.n_steps},
                                   self.n_steps, self.thinning, outputs)
