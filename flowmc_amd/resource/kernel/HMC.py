"""HMC proposal — src/flowMC/resource/kernel/HMC.py:10-163."""

from __future__ import annotations

import ctypes as C

import torch

from ... import _native
from ... import random as frandom
from ..logPDF import LogPDF
from .base import KERNEL_HMC, ProposalBase, StepOutputs, steps_needed
from .MALA import _prep_state


def condition_spec(M, n_dim: int, device):
    """Kernel form of a condition matrix M (HMC.py:118-136): (kinetic
    weights [d] or None for the identity, momentum factor [d, d] or None).
    A scalar c means c * I and a vector its diagonal; a diagonal matrix gives
    its diagonal; a dense matrix gives (colsum(M), chol(inv(M))), both
    formed in FP64 on the host and stored in FP32."""
    if isinstance(M, (int, float)) or torch.as_tensor(M).ndim == 0:
        c = float(M)
        return (None, None) if c == 1.0 else (torch.full((n_dim,), c, device=device, dtype=torch.float32), None)
    Mt = torch.as_tensor(M, dtype=torch.float64).cpu()
    if Mt.ndim == 1:
        if Mt.numel() != n_dim:
            raise ValueError(f"condition_matrix must have {n_dim} diagonal entries, got {Mt.numel()}")
        return Mt.to(device, torch.float32).contiguous(), None
    if Mt.ndim != 2:
        raise ValueError("condition_matrix must be a scalar, a vector or a matrix")
    if tuple(Mt.shape) != (n_dim, n_dim):
        raise ValueError(f"condition_matrix must be [{n_dim}, {n_dim}], got {list(Mt.shape)}")
    off = Mt - torch.diag(torch.diagonal(Mt))
    if torch.count_nonzero(off) == 0:
        diag = torch.diagonal(Mt)
        return (None, None) if bool(torch.all(diag == 1)) else (diag.to(device, torch.float32).contiguous(), None)
    chol = torch.linalg.cholesky(torch.linalg.inv(Mt))
    return Mt.sum(0).to(device, torch.float32).contiguous(), chol.to(device, torch.float32).contiguous()


class HMC(ProposalBase):
    """Hamiltonian Monte Carlo with the reference's leapfrog schedule
    (n_leapfrog + 1 drifts, n_leapfrog + 2 gradient kicks, HMC.py:47-50,91-95).

    ``condition_matrix`` may be a scalar c (c * I), a vector (its diagonal)
    or a [d, d] matrix. The momentum is xi @ chol(inv(M)).T (HMC.py:133-136),
    xi / sqrt(m) for a diagonal M, and the kinetic energy the reference's
    broadcast 1/2 sum(p^2 * M) (HMC.py:118-121): 1/2 sum_j m_j p_j^2 for a
    diagonal M, 1/2 sum_j colsum_j(M) p_j^2 for a dense one (whose
    p-gradient colsum(M) * p is then the drift).
    """

    step_size: float

    @property
    def n_leapfrog(self) -> int:
        return self._n_leapfrog

    def __repr__(self):
        return "HMC with step size " + str(self.step_size) + " and " + str(self.n_leapfrog) + " leapfrog steps"

    def __init__(self, condition_matrix=1, step_size: float = 0.1, n_leapfrog: int = 10, fused_steps: int = 16,
                 launch_cfg=None):
        self.condition_matrix = condition_matrix
        self.step_size = step_size
        self._n_leapfrog = int(n_leapfrog)
        self.fused_steps = int(fused_steps)
        self.launch_cfg = launch_cfg
        coefs = torch.ones(self._n_leapfrog + 2, 2, dtype=torch.float32)
        coefs[0] = torch.tensor([0.0, 0.5])
        coefs[-1] = torch.tensor([1.0, 0.5])
        self.leapfrog_coefs = coefs  # HMC.py:46-49 table (documentation / tests)

    def metric(self, n_dim: int, device) -> torch.Tensor | None:
        """Kinetic-energy weights as a device vector (None = identity): the
        diagonal of a diagonal condition matrix, colsum(M) of a dense one."""
        return self.metric_spec(n_dim, device)[0]

    def metric_spec(self, n_dim: int, device):
        """(kinetic weights [d] or None, momentum factor chol(inv(M)) [d, d]
        or None) of the condition matrix (see ``condition_spec``)."""
        key = (int(n_dim), str(device), id(self.condition_matrix))
        cache = self.__dict__.setdefault("_metric_cache", {})
        if key not in cache:
            cache.clear()
            cache[key] = condition_spec(self.condition_matrix, n_dim, device)
        return cache[key]

    def kernel(self, rng_key, position, log_prob, logpdf, data):
        """One HMC step (HMC.py:98-151); H0 uses the passed-in ``log_prob``."""
        single = torch.as_tensor(position).ndim == 1
        x = _prep_state(position)
        x = x[None] if single else x
        C_, d = x.shape
        lp_in = torch.as_tensor(log_prob, dtype=torch.float32, device=x.device).reshape(-1).expand(C_).contiguous()
        lp_out = torch.empty(C_, 1, device=x.device, dtype=torch.float32)
        acc = torch.empty(C_, 1, device=x.device, dtype=torch.float32)
        out = StepOutputs(None, lp_out, acc, 0)
        self._run_generic(logpdf, data, x, 1, *frandom.key_words(rng_key), None, None, out, 1, lp_in=lp_in)
        if single:
            return x[0], lp_out[0, 0], acc[0, 0] > 0
        return x, lp_out[:, 0], acc[:, 0] > 0

    def sample_steps(self, rng_key, initial_position, logpdf, data, n_steps, thinning, outputs: StepOutputs,
                     noise=None, log_u=None):
        x = _prep_state(initial_position)
        n_run = steps_needed(n_steps, thinning)
        k0, k1 = frandom.key_words(rng_key)
        if noise is not None:
            noise = noise.to(torch.float32).contiguous()
            log_u = log_u.to(torch.float32).contiguous()
        if isinstance(logpdf, LogPDF) and logpdf.is_native:
            self._run_fused(logpdf.log_pdf, x, n_run, k0, k1, noise, log_u, outputs, thinning)
        else:
            self._run_generic(logpdf, data, x, n_run, k0, k1, noise, log_u, outputs, thinning)
        return x

    def _run_fused(self, target, x, n_run, k0, k1, noise, log_u, out: StepOutputs, thinning):
        C_, d = x.shape
        desc = target.descriptor()
        m, chol = self.metric_spec(d, x.device)
        lp = torch.empty(C_, device=x.device, dtype=torch.float32)
        st = _native.stream_ptr()
        cfg = C.byref(self.launch_cfg) if self.launch_cfg is not None else None
        K = max(1, self.fused_steps)
        for s0 in range(0, n_run, K):
            ns = min(K, n_run - s0)
            nz = noise[s0:s0 + ns] if noise is not None else None
            lu = log_u[s0:s0 + ns] if log_u is not None else None
            _native.call(
                "fmc_hmc_run", st, C.byref(desc), _native.dptr(x), _native.dptr(lp), C_, ns, s0,
                float(self.step_size), self.n_leapfrog, _native.dptr(m, name="metric", numel=d if m is not None else None),
                _native.dptr(chol, name="momentum factor", numel=d * d if chol is not None else None), k0, k1,
                _native.dptr(nz, name="noise"), _native.dptr(lu, name="log_u"),
                _native.dptr(out.positions, name="position buffer"), _native.dptr(out.log_probs, name="log_prob buffer"),
                _native.dptr(out.accepts, name="acceptance buffer"), out.buf_len, out.cursor, thinning, cfg)
        self.last_log_prob = lp

    def _run_generic(self, logpdf, data, x, n_run, k0, k1, noise, log_u, out: StepOutputs, thinning, lp_in=None):
        C_, d = x.shape
        st = _native.stream_ptr()
        m, chol = self.metric_spec(d, x.device)
        mp = _native.dptr(m, name="metric", numel=d if m is not None else None)
        cp = _native.dptr(chol, name="momentum factor", numel=d * d if chol is not None else None)
        lp, g = logpdf.value_and_grad(x, data)
        lp, g = lp.contiguous().clone(), g.contiguous().clone()
        if lp_in is not None:
            lp = lp_in.clone()
        p = torch.empty_like(x)
        k_0 = torch.empty(C_, dtype=torch.float64, device=x.device)
        eps = float(self.step_size)
        L = self.n_leapfrog
        for s in range(n_run):
            nz = noise[s].contiguous() if noise is not None else None
            _native.call("fmc_hmc_momentum", st, _native.dptr(p), _native.dptr(k_0, torch.float64), C_, d, mp, cp,
                         k0, k1, s, _native.dptr(nz, name="noise"))
            xl = x.clone()
            _native.call("fmc_hmc_kick", st, _native.dptr(p), _native.dptr(g), C_, d, eps * 0.5)
            gl = g
            for i in range(1, L + 2):
                _native.call("fmc_hmc_drift", st, _native.dptr(xl), _native.dptr(p), mp, C_, d, eps)
                lpl, gl = logpdf.value_and_grad(xl, data)
                gl = gl.contiguous()
                _native.call("fmc_hmc_kick", st, _native.dptr(p), _native.dptr(gl), C_, d,
                             eps * (1.0 if i <= L else 0.5))
            lu = log_u[s].contiguous() if log_u is not None else None
            lpl = lpl.contiguous()
            _native.call("fmc_hmc_accept", st, _native.dptr(x), _native.dptr(lp), _native.dptr(xl),
                         _native.dptr(lpl), _native.dptr(p), _native.dptr(k_0, torch.float64), mp, C_, d,
                         k0, k1, s, _native.dptr(lu, name="log_u"),
                         _native.dptr(out.positions, name="position buffer"),
                         _native.dptr(out.log_probs, name="log_prob buffer"),
                         _native.dptr(out.accepts, name="acceptance buffer"), out.buf_len, out.cursor, thinning)
            # the gradient at the (possibly new) state: equal positions have equal gradients
            same = (x == xl).all(-1, keepdim=True)
            g = torch.where(same, gl, g).contiguous()
        self.last_log_prob = lp

    def _tempered_launch(self, n_dim, device):
        return (KERNEL_HMC, self.step_size, self.n_leapfrog, *self.metric_spec(n_dim, device))

    def print_parameters(self):
        print("HMC parameters:")
        print(f"step_size: {self.step_size}")
        print(f"n_leapfrog: {self.n_leapfrog}")
        print(f"condition_matrix: {self.condition_matrix}")

    def save_resource(self, path):
        raise NotImplementedError

    def load_resource(self, path):
        raise NotImplementedError
