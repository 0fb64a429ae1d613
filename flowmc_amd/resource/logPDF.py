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


class TemperedPDF(LogPDF):
    """Tempered log-density — src/flowMC/resource/logPDF.py:85-115.

    ``tempered_log_pdf(T, x, data) = log_likelihood(x, data) / T + log_prior(x, data)``;
    calling the object returns the untempered log-likelihood (logPDF.py:101-102),
    which is what the exchange step compares. ``n_temps`` and ``max_temp`` are
    accepted and unused, as in the reference.

    ``log_prior=None`` is the zero prior: with a built-in target the tempered
    local steps then run in the fused chain kernel (fmc_tempered_run); any other
    prior is evaluated with torch and the steps take the per-step path.
    """

    def __repr__(self):
        return "TemperedPDF with " + str(self.n_dims) + " dimensions"

    def __init__(self, log_likelihood: Callable, log_prior: Optional[Callable] = None, variables=None,
                 n_dims: Optional[int] = None, n_temps: int = 5, max_temp: float = 100, batched: bool = False):
        if isinstance(log_likelihood, LogPDF):
            batched = log_likelihood.batched
            if variables is None and n_dims is None:
                variables = log_likelihood.variables
            log_likelihood = log_likelihood.log_pdf
        super().__init__(log_likelihood, variables, n_dims, batched=batched)
        self.log_prior = log_prior

    @property
    def fused_tempering(self) -> bool:
        """Built-in likelihood and zero prior: tempered steps run fused on the device."""
        return self.is_native and self.log_prior is None

    def _prior(self, x: torch.Tensor, data) -> torch.Tensor:
        if self.log_prior is None:
            return torch.zeros(x.shape[:-1], dtype=torch.float32, device=x.device)
        if x.ndim == 1:
            return torch.as_tensor(self.log_prior(x, data), dtype=torch.float32, device=x.device).reshape(())
        flat = x.reshape(-1, x.shape[-1])
        try:
            v = torch.func.vmap(self.log_prior, in_dims=(0, None))(flat, data)
        except (RuntimeError, ValueError):
            v = torch.stack([torch.as_tensor(self.log_prior(xi, data)).reshape(()) for xi in flat])
        return torch.as_tensor(v, dtype=torch.float32, device=x.device).reshape(x.shape[:-1])

    def tempered_log_pdf(self, temperatures, x: torch.Tensor, data=None) -> torch.Tensor:
        """(1 / T) * log_likelihood(x) + log_prior(x); T broadcasts against x's leading dims."""
        base = self(x, data)
        T = torch.as_tensor(temperatures, dtype=base.dtype, device=base.device)
        return (1.0 / T) * base + self._prior(x, data)


class TemperedRows(LogPDF):
    """Rows [R, d] of tempered chains seen as one batched log-density: row r
    targets tempered_log_pdf(temperatures[r % n_temps], x_r). The per-step
    kernel paths evaluate and differentiate it like any batched LogPDF."""

    def __init__(self, tempered: TemperedPDF, temperatures: torch.Tensor):
        super().__init__(tempered.log_pdf, tempered.variables, batched=True)
        self.tempered = tempered
        self.temperatures = torch.as_tensor(temperatures, dtype=torch.float32)

    @property
    def is_native(self) -> bool:
        return False

    def _row_temps(self, x):
        T = self.temperatures.to(x.device)
        return T.repeat(x.shape[0] // T.numel())

    def __call__(self, x: torch.Tensor, data=None) -> torch.Tensor:
        return self.tempered.tempered_log_pdf(self._row_temps(x), x, data)

    def value_and_grad(self, x: torch.Tensor, data=None):
        t = self.tempered
        beta = 1.0 / self._row_temps(x)
        if t.is_native:
            v, g = t.log_pdf.value_and_grad(x)
        else:
            v, g = LogPDF.value_and_grad(t, x, data)
        v, g = beta * v, beta[:, None] * g
        if t.log_prior is not None:
            with torch.enable_grad():
                xr = x.detach().clone().requires_grad_(True)
                pv = t._prior(xr, data)
                if pv.requires_grad:
                    (pg,) = torch.autograd.grad(pv.sum(), xr)
                else:
                    pg = torch.zeros_like(xr)
            v, g = v + pv.detach(), g + pg
        return v.to(torch.float32).contiguous(), g.to(torch.float32).contiguous()
