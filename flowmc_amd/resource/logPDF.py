"""LogPDF resource — src/flowMC/resource/logPDF.py:20-81.

``log_pdf`` is either a built-in :class:`flowmc_amd.targets.Target` (evaluated
by the fused device kernels with a hand-written gradient) or a Python callable
``f(x[d], data) -> scalar`` written with torch ops; the latter is batched over
chains with ``torch.func.vmap`` and differentiated with ``torch.func.grad``.
A callable that already maps [N, d] -> [N] (``batched=True``, or detected when
vmap cannot trace it, e.g. because it calls device kernels such as a flow's
``log_prob``) is called on the whole batch and differentiated with
torch.autograd.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Optional

import torch

from ..targets import Target
from .base import Resource


@dataclass
class Variable:
    name: str
    continuous: bool


class LogPDF(Resource):
    log_pdf: Callable
    variables: list

    @property
    def n_dims(self):
        return len(self.variables)

    @property
    def is_native(self) -> bool:
        return isinstance(self.log_pdf, Target)

    def __repr__(self):
        return "LogPDF with " + str(self.n_dims) + " dimensions"

    def __init__(self, log_pdf: Callable, variables: Optional[list] = None, n_dims: Optional[int] = None,
                 batched: bool = False):
        self.log_pdf = log_pdf
        self.batched = batched
        if variables is None and n_dims is not None:
            self.variables = [Variable("x_" + str(i), True) for i in range(n_dims)]
        elif variables is not None:
            self.variables = variables
        else:
            raise ValueError("Either variables or n_dims must be provided")
        self._vg = None

    def __call__(self, x: torch.Tensor, data=None) -> torch.Tensor:
        """Log-density of one point [d] or a batch [..., d]."""
        if self.is_native:
            return self.log_pdf(x, data)
        if x.ndim == 1:
            return self.log_pdf(x[None], data)[0] if self.batched else self.log_pdf(x, data)
        flat = x.reshape(-1, x.shape[-1])
        return self._batch_call(flat, data).reshape(x.shape[:-1])

    def _batch_call(self, flat, data):
        if not self.batched:
            try:
                return torch.func.vmap(self.log_pdf, in_dims=(0, None))(flat, data)
            except RuntimeError:
                out = self.log_pdf(flat, data)
                if tuple(out.shape) != (flat.shape[0],):
                    raise
                self.batched = True
                return out
        return self.log_pdf(flat, data)

    def value_and_grad(self, x: torch.Tensor, data=None):
        """Batched (log p, d log p / dx) at x [C, d]."""
        if self.is_native:
            return self.log_pdf.value_and_grad(x)
        if not self.batched:
            try:
                self._batch_call(x[:1], data)
            except RuntimeError:
                pass
        if self.batched:
            with torch.enable_grad():
                xr = x.detach().clone().requires_grad_(True)
                v = self.log_pdf(xr, data)
                (g,) = torch.autograd.grad(v.sum(), xr)
            return v.detach().to(torch.float32).contiguous(), g.to(torch.float32).contiguous()
        if self._vg is None:
            def vg(xi, d):
                g, v = torch.func.grad_and_value(self.log_pdf)(xi, d)
                return v, g
            self._vg = torch.func.vmap(vg, in_dims=(0, None))
        v, g = self._vg(x, data)
        return v.to(torch.float32).contiguous(), g.to(torch.float32).contiguous()

    def print_parameters(self):
        print("LogPDF with variables:")
        for var in self.variables:
            print(var.name, var.continuous)

    def save_resource(self, path):
        raise NotImplementedError

    def load_resource(self, path):
        raise NotImplementedError
