"""ParallelTempering strategy — src/flowMC/strategy/parallel_tempering.py:13-436.

One call: ``n_steps`` local steps of every (chain, temperature) walker on the
tempered target, one sweep of adjacent-temperature exchanges, and, while the
sampler state says "training", the tempered walkers and the adapted
temperatures are written back (parallel_tempering.py:116-127). Only the
temperature-1 walkers are returned; in production the tempered buffer is left
as training last wrote it, so every production call restarts the tempered
walkers from there, as in the reference.

The reference vmaps per-chain helpers over chains; here each helper works on
all chains at once:

* ``_ensemble_step`` -> ``ProposalBase.sample_tempered``: one fused chain
  launch over C*T rows for built-in targets (fmc_tempered_run), or the
  per-step kernels around a torch-evaluated tempered density otherwise;
* ``_exchange`` -> fmc_pt_exchange (per-chain sequential swap scan + gather),
  on the untempered log-likelihoods of all walkers;
* ``_adapt_temperature`` -> fmc_pt_adapt (one block, in place).
"""

from __future__ import annotations

import torch

from .. import _native
from .. import random as frandom
from ..resource.buffers import Buffer
from ..resource.kernel.base import ProposalBase
from ..resource.kernel.NF_proposal import _eval_logpdf
from ..resource.logPDF import TemperedPDF
from ..resource.states import State
from .base import Strategy


class ParallelTempering(Strategy):
    n_steps: int
    tempered_logpdf_name: str
    kernel_name: str
    tempered_buffer_names: list
    verbose: bool = False

    def __init__(self, n_steps: int, tempered_logpdf_name: str, kernel_name: str, tempered_buffer_names: list,
                 state_name: str, verbose: bool = False):
        self.n_steps = n_steps
        self.tempered_logpdf_name = tempered_logpdf_name
        self.kernel_name = kernel_name
        self.tempered_buffer_names = tempered_buffer_names
        self.verbose = verbose
        self.state_name = state_name

    def __call__(self, rng_key, resources, initial_position, data):
        rng_key, subkey = frandom.split(rng_key)
        assert isinstance(kernel := resources[self.kernel_name], ProposalBase)
        assert isinstance(tempered_logpdf := resources[self.tempered_logpdf_name], TemperedPDF)
        assert isinstance(tempered_positions := resources[self.tempered_buffer_names[0]], Buffer)
        assert isinstance(temperatures := resources[self.tempered_buffer_names[1]], Buffer)
        assert isinstance(state := resources[self.state_name], State)
        _native.require_gpu()
        x0 = torch.as_tensor(initial_position, dtype=torch.float32)
        x0 = x0 if x0.is_cuda else x0.cuda()
        positions = torch.cat([x0[:, None, :], tempered_positions.data.to(x0.device)], dim=1).contiguous()
        temps = temperatures.data.to(device=x0.device, dtype=torch.float32).contiguous()

        rng_key, subkey = frandom.split(rng_key)
        positions, log_probs, do_accepts = self._ensemble_step(kernel, subkey, positions, tempered_logpdf, temps,
                                                               data)
        if self.verbose:
            print("Mean acceptance of individual steps in PT: " + str(float(do_accepts.mean())))

        rng_key, subkey = frandom.split(rng_key)
        positions, log_probs, do_accepts = self._exchange(subkey, positions, tempered_logpdf, temps, data)
        if self.verbose:
            print("Mean acceptance of exchange steps in PT: " + str(float(do_accepts.mean())))

        if state.data["training"]:
            tempered_positions.update_buffer(positions[:, 1:], 0)
            temperatures.update_buffer(self._adapt_temperature(temps, do_accepts), 0)
        return rng_key, resources, positions[:, 0]

    # ------------------------------------------------------------------
    def _individual_step_body(self, kernel: ProposalBase, carry, aux):
        """One tempered step of one walker (parallel_tempering.py:135-205):
        carry = (key, position [d], log_prob, logpdf, temperature, data);
        returns the new carry and (position, log_prob, do_accept)."""
        key, position, log_prob, logpdf, temperature, data = carry
        key, subkey = frandom.split(key)
        pos, lp, acc = self._walker_steps(kernel, subkey, position, logpdf, temperature, data, 1)
        return (key, pos, lp, logpdf, temperature, data), (pos, lp, acc[0])

    def _individal_step(self, kernel: ProposalBase, rng_key, positions, logpdf: TemperedPDF, temperatures, data):
        """n_steps tempered steps of one walker (parallel_tempering.py:207-248,
        the reference's spelling): returns the last position [d], its tempered
        log-probability and the step decisions [n_steps]."""
        return self._walker_steps(kernel, rng_key, positions, logpdf, temperatures, data, self.n_steps)

    def _walker_steps(self, kernel, key, position, logpdf, temperature, data, n):
        x = torch.as_tensor(position, dtype=torch.float32)
        x = (x if x.is_cuda else x.cuda()).reshape(1, 1, -1)
        t = torch.as_tensor(temperature, dtype=torch.float32).reshape(1).to(x.device)
        acc = torch.empty(1, max(1, n), dtype=torch.float32, device=x.device)
        pos, lp = kernel.sample_tempered(key, x, logpdf, t, data, n, accepts=acc)
        return pos.reshape(-1), lp.reshape(()), acc[0, :n] > 0

    def _ensemble_step(self, kernel: ProposalBase, rng_key, positions: torch.Tensor, logpdf: TemperedPDF,
                       temperatures: torch.Tensor, data, noise=None, log_u=None):
        """n_steps local steps of all walkers (parallel_tempering.py:
        135-290): positions [C, T, d], or [T, d] for one chain as in the
        reference's per-chain helper. Returns positions, tempered log-probs
        [(C,) T] and the step decisions [(C,) T, n_steps]."""
        x = torch.as_tensor(positions, dtype=torch.float32)
        x = x if x.is_cuda else x.cuda()
        single = x.ndim == 2
        x = x[None] if single else x
        C_, T, _ = x.shape
        temps = torch.as_tensor(temperatures, dtype=torch.float32).to(x.device)
        acc = torch.empty(C_ * T, max(1, self.n_steps), dtype=torch.float32, device=temps.device)
        pos, lp = kernel.sample_tempered(rng_key, x, logpdf, temps, data, self.n_steps, noise=noise,
                                         log_u=log_u, accepts=acc)
        acc = acc[:, :self.n_steps].reshape(C_, T, self.n_steps) > 0
        return (pos[0], lp[0], acc[0]) if single else (pos, lp, acc)

    def _exchange(self, rng_key, positions: torch.Tensor, logpdf: TemperedPDF, temperatures: torch.Tensor, data,
                  log_u=None):
        """One sweep of adjacent swaps per chain (parallel_tempering.py:291-398)
        on the untempered log-likelihoods (``logpdf(x, data)``). positions
        [C, T, d], or [T, d] for one chain. Returns the positions, their
        log-likelihoods [(C,) T] and the swap decisions [(C,) T-1]."""
        single = positions.ndim == 2
        x = torch.as_tensor(positions, dtype=torch.float32)
        x = (x if x.is_cuda else x.cuda())
        x = (x[None] if single else x).contiguous()
        C_, T, d = x.shape
        temps = torch.as_tensor(temperatures, dtype=torch.float32).to(x.device).contiguous()
        loglik = _eval_logpdf(logpdf, x.reshape(C_ * T, d), data).reshape(C_, T).to(torch.float32).contiguous()
        loglik = loglik.clone()
        out = torch.empty_like(x)
        acc = torch.empty(C_, max(T - 1, 1), dtype=torch.float32, device=x.device)
        k0, k1 = frandom.key_words(rng_key)
        lu = log_u.to(device=x.device, dtype=torch.float32).contiguous() if log_u is not None else None
        _native.call("fmc_pt_exchange", _native.stream_ptr(), _native.dptr(x, name="positions"),
                     _native.dptr(out), _native.dptr(loglik), _native.dptr(temps, name="temperatures", numel=T),
                     C_, T, d, k0, k1, _native.dptr(lu, name="log_u", numel=C_ * (T - 1) if lu is not None else None),
                     _native.dptr(acc))
        acc = acc[:, :T - 1] > 0
        return (out[0], loglik[0], acc[0]) if single else (out, loglik, acc)

    def _adapt_temperature(self, temperatures: torch.Tensor, do_accept: torch.Tensor) -> torch.Tensor:
        """Temperature adaptation (parallel_tempering.py:400-436) from the swap
        decisions [C, >= T-1]; returns the new temperatures (the input is not
        modified). As in the reference, only the first T-1 columns matter."""
        temps = torch.as_tensor(temperatures, dtype=torch.float32)
        temps = (temps if temps.is_cuda else temps.cuda()).contiguous().clone()
        T = temps.numel()
        acc = torch.as_tensor(do_accept).to(device=temps.device, dtype=torch.float32)
        C_ = acc.shape[0]
        if T >= 3:
            if acc.shape[1] < T - 1:
                raise ValueError(f"need swap decisions for {T - 1} pairs, got {acc.shape[1]}")
            acc = acc[:, :T - 1].contiguous()
            _native.call("fmc_pt_adapt", _native.stream_ptr(), _native.dptr(temps), _native.dptr(acc), C_, T)
        return temps
