"""MALA proposal — src/flowMC/resource/kernel/MALA.py:11-99."""

from __future__ import annotations

import ctypes as C

import torch

from ... import _native
from ... import random as frandom
from ..logPDF import LogPDF
from .base import KERNEL_MALA, ProposalBase, StepOutputs, steps_needed


def _prep_state(position: torch.Tensor) -> torch.Tensor:
    _native.require_gpu()
    x = torch.as_tensor(position, dtype=torch.float32)
    if not x.is_cuda:
        x = x.cuda()
    return x.contiguous().clone()


class MALA(ProposalBase):
    """Metropolis-adjusted Langevin algorithm.

    Built-in targets run ``fused_steps`` steps per launch of the fused chain
    kernel (fmc_mala_run); other targets run one propose / accept kernel pair
    per step around a torch.func gradient.
    """

    step_size: float

    def __repr__(self):
        return "MALA with step size " + str(self.step_size)

    def __init__(self, step_size: float, fused_steps: int = 64, launch_cfg=None):
        super().__init__()
        self.step_size = step_size
        self.fused_steps = int(fused_steps)
        self.launch_cfg = launch_cfg  # optional _native.fmc_launch_cfg (autotuner output)

    # ----------------------------------------------------------------------
    def kernel(self, rng_key, position, log_prob, logpdf, data):
        """One MALA step (MALA.py:26-89). ``log_prob`` is ignored, as in the
        reference, which re-evaluates the target at ``position``."""
        single = torch.as_tensor(position).ndim == 1
        x = _prep_state(position)
        x = x[None] if single else x
        C_, d = x.shape
        lp = torch.empty(C_, 1, device=x.device, dtype=torch.float32)
        acc = torch.empty(C_, 1, device=x.device, dtype=torch.float32)
        out = StepOutputs(None, lp, acc, 0)
        self.sample_steps(rng_key, x, logpdf, data, 1, 1, out, _inplace=True)
        x, lp, acc = x, lp[:, 0], acc[:, 0] > 0
        if single:
            return x[0], lp[0], acc[0]
        return x, lp, acc

    # ----------------------------------------------------------------------
    def sample_steps(self, rng_key, initial_position, logpdf, data, n_steps, thinning, outputs: StepOutputs,
                     noise=None, log_u=None, _inplace=False):
        x = initial_position if _inplace else _prep_state(initial_position)
        C_, d = x.shape
        n_run = steps_needed(n_steps, thinning)
        k0, k1 = frandom.key_words(rng_key)
        if noise is not None:
            noise = noise.to(torch.float32).contiguous()
            log_u = log_u.to(torch.float32).contiguous()
            if noise.shape[0] < n_run or log_u.shape[0] < n_run:
                raise ValueError("injected noise must cover every step that is run")
        if isinstance(logpdf, LogPDF) and logpdf.is_native:
            self._run_fused(logpdf.log_pdf, x, n_run, k0, k1, noise, log_u, outputs, thinning)
        else:
            self._run_generic(logpdf, data, x, n_run, k0, k1, noise, log_u, outputs, thinning)
        return x

    def _run_fused(self, target, x, n_run, k0, k1, noise, log_u, out: StepOutputs, thinning):
        C_, d = x.shape
        desc = target.descriptor()
        lp = torch.empty(C_, device=x.device, dtype=torch.float32)
        st = _native.stream_ptr()
        cfg = C.byref(self.launch_cfg) if self.launch_cfg is not None else None
        K = max(1, self.fused_steps)
        for s0 in range(0, n_run, K):
            ns = min(K, n_run - s0)
            nz = noise[s0:s0 + ns] if noise is not None else None
            lu = log_u[s0:s0 + ns] if log_u is not None else None
            _native.call(
                "fmc_mala_run", st, C.byref(desc), _native.dptr(x, name="x"), _native.dptr(lp, name="logp"),
                C_, ns, s0, float(self.step_size), k0, k1,
                _native.dptr(nz, name="noise") if nz is not None else None,
                _native.dptr(lu, name="log_u") if lu is not None else None,
                _native.dptr(out.positions, name="position buffer"),
                _native.dptr(out.log_probs, name="log_prob buffer"),
                _native.dptr(out.accepts, name="acceptance buffer"),
                out.buf_len, out.cursor, thinning, cfg)
        self.last_log_prob = lp

    def _run_generic(self, logpdf, data, x, n_run, k0, k1, noise, log_u, out: StepOutputs, thinning):
        C_, d = x.shape
        st = _native.stream_ptr()
        lp, g = logpdf.value_and_grad(x, data)
        lp, g = lp.contiguous().clone(), g.contiguous().clone()
        prop = torch.empty_like(x)
        for s in range(n_run):
            nz = noise[s].contiguous() if noise is not None else None
            _native.call("fmc_mala_propose", st, _native.dptr(x), _native.dptr(g), _native.dptr(prop), C_, d,
                         float(self.step_size), k0, k1, s, _native.dptr(nz, name="noise"))
            lp1, g1 = logpdf.value_and_grad(prop, data)
            lp1, g1 = lp1.contiguous(), g1.contiguous()  # bound names keep them alive across the call
            lu = log_u[s].contiguous() if log_u is not None else None
            _native.call("fmc_mala_accept", st, _native.dptr(x), _native.dptr(lp), _native.dptr(g),
                         _native.dptr(prop), _native.dptr(lp1), _native.dptr(g1),
                         C_, d, float(self.step_size), k0, k1, s, _native.dptr(lu, name="log_u"),
                         _native.dptr(out.positions, name="position buffer"),
                         _native.dptr(out.log_probs, name="log_prob buffer"),
                         _native.dptr(out.accepts, name="acceptance buffer"), out.buf_len, out.cursor, thinning)
        self.last_log_prob = lp

    def _tempered_launch(self, n_dim, device):
        return KERNEL_MALA, self.step_size, 0, None

    def print_parameters(self):
        print("MALA parameters:")
        print(f"step_size: {self.step_size}")

    def save_resource(self, path):
        raise NotImplementedError

    def load_resource(self, path):
        raise NotImplementedError
