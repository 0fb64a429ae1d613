"""NFProposal — src/flowMC/resource/kernel/NF_proposal.py:15-184.

Per chain, ``n_steps`` proposals are drawn from the flow (inverse pass over
Philox base noise, one launch for all C*S proposals), their flow log-density
is evaluated by a forward pass (the reference evaluates log_prob explicitly,
NF_proposal.py:150-167, so the inverse-order quirk of the flow is honoured);
with ``exact_inverse`` it comes from the sampling pass itself,
the target is evaluated on them, and the independence Metropolis–Hastings
scan (NF_proposal.py:91-126) runs in fmc_nf_mh, which writes the thinned
buffer rows directly.

Base-noise counter of chain c, proposal s: sample index c*S + s; MH uniform
counter: (0, c, s, stream 4).
"""

from __future__ import annotations

import ctypes as C

import torch

from ... import _native
from ... import random as frandom
from .base import ProposalBase, StepOutputs, rows_written


class NFProposal(ProposalBase):
    def __repr__(self):
        return "NF proposal with " + self.model.__repr__()

    def __init__(self, model, n_NFproposal_batch_size: int = 100):
        super().__init__()
        self.model = model
        # Proposals are generated in one launch per pass; the reference's
        # micro-batching (NF_proposal.py:50-89) only bounds JAX memory and
        # does not change results, so the value is kept for API compatibility.
        self.n_batch_size = n_NFproposal_batch_size

    def kernel(self, rng_key, position, log_prob, logpdf, data):
        """NF_proposal.py:27-128 for one chain ([d]) or a batch ([C, d]); data["n_steps"] proposals."""
        n_steps = int(data["n_steps"])
        single = torch.as_tensor(position).ndim == 1
        x = torch.as_tensor(position, dtype=torch.float32).cuda().contiguous()
        x = x[None] if single else x
        C_, d = x.shape
        pos = torch.empty((C_, n_steps, d), dtype=torch.float32, device=x.device)
        lp = torch.empty((C_, n_steps), dtype=torch.float32, device=x.device)
        acc = torch.empty((C_, n_steps), dtype=torch.float32, device=x.device)
        self.sample_steps(rng_key, x, logpdf, data, n_steps, 1, StepOutputs(pos, lp, acc, 0))
        if single:
            return pos[0], lp[0], acc[0] > 0
        return pos, lp, acc > 0

    def sample_steps(self, rng_key, initial_position, logpdf, data, n_steps, thinning, outputs: StepOutputs,
                     noise=None, log_u=None):
        _native.require_gpu()
        x0 = torch.as_tensor(initial_position, dtype=torch.float32).cuda().contiguous()
        C_, d = x0.shape
        S = int(n_steps)
        model = self.model
        k0, k1 = frandom.key_words(rng_key)
        # 1. proposals [C*S, d] from the flow and 2. their flow log-density.
        # With the exact inverse one pass gives both (log N(z) minus the
        # inverse log-det); the reference's layer order needs the explicit
        # forward log_prob (NF_proposal.py:150-167).
        if getattr(model, "exact_inverse", False):
            if noise is not None:
                prop, q_prop = model.sample_and_log_prob(noise=noise.reshape(C_ * S, d))
            else:
                prop, q_prop = model.sample_and_log_prob(rng_key, C_ * S)
        else:
            if noise is not None:
                prop = model.sample_from_noise(noise.reshape(C_ * S, d))
            else:
                prop = model.sample(rng_key, C_ * S)
            q_prop = model.log_prob(prop)
        q0 = model.log_prob(x0)
        # 3. target log-density (take_steps.py:202 evaluates the current point)
        lp_prop = _eval_logpdf(logpdf, prop, data)
        lp0 = _eval_logpdf(logpdf, x0, data)
        idx = torch.empty((C_, S), dtype=torch.int32, device=x0.device)
        x_last = torch.empty_like(x0)
        lp_last = torch.empty(C_, dtype=torch.float32, device=x0.device)
        lu = log_u.to(torch.float32).reshape(C_, S).contiguous() if log_u is not None else None
        _native.call("fmc_nf_mh", _native.stream_ptr(), x0.data_ptr(), lp0.data_ptr(), q0.data_ptr(),
                     prop.data_ptr(), lp_prop.data_ptr(), q_prop.data_ptr(), C_, S, d, k0, k1,
                     _native.dptr(lu, name="log_u"), idx.data_ptr(), x_last.data_ptr(), lp_last.data_ptr(),
                     _native.dptr(outputs.positions, name="position buffer"),
                     _native.dptr(outputs.log_probs, name="log_prob buffer"),
                     _native.dptr(outputs.accepts, name="acceptance buffer"), outputs.buf_len, outputs.cursor,
                     thinning)
        self.last_log_prob = lp_last
        return x_last

    def print_parameters(self):
        raise NotImplementedError

    def save_resource(self, path):
        raise NotImplementedError

    def load_resource(self, path):
        raise NotImplementedError


def _eval_logpdf(logpdf, x, data):
    from ..logPDF import LogPDF

    if isinstance(logpdf, LogPDF) and logpdf.is_native:
        lp, _ = logpdf.log_pdf.value_and_grad(x, need_grad=False)
        return lp.contiguous()
    with torch.no_grad():
        return logpdf(x, data).to(torch.float32).contiguous()
