"""FP64 restatement of parallel tempering — src/flowMC/strategy/parallel_tempering.py.

Test infrastructure only (see the package docstring). Walkers are
[C, T, d] (chain, temperature slot, dimension); randomness is injected.
"""

from __future__ import annotations

import torch

from .chains import take_serial_steps


def tempered_rows(fn, temperatures, n_chains):
    """Row c*T + t of a [C*T, d] batch targets fn(x) / T_t —
    TemperedPDF.tempered_log_pdf (src/flowMC/resource/logPDF.py:104-106) with
    a zero log prior."""
    T = torch.as_tensor(temperatures, dtype=torch.float64)
    beta = (1.0 / T).repeat(n_chains)

    def f(x):
        return beta * fn(x)

    return f


def ensemble_steps(kind, x, fn, temperatures, noise, log_u, **kw):
    """ParallelTempering._ensemble_step (parallel_tempering.py:250-290) for all
    chains: each walker takes noise.shape[0] local steps on its tempered
    target. noise [S, C*T, d], log_u [S, C*T]. Returns positions [C, T, d],
    tempered log-probs [C, T], decisions [C, T, S] and log-ratios [C*T, S]."""
    C, T, d = x.shape
    rows = x.reshape(C * T, d)
    pos, lps, accs, last, ratios = take_serial_steps(kind, rows, tempered_rows(fn, temperatures, C), noise, log_u,
                                                     1, **kw)
    return last.reshape(C, T, d), lps[:, -1].reshape(C, T), accs.reshape(C, T, -1), ratios


def exchange(x, loglik, temperatures, log_u):
    """ParallelTempering._exchange (parallel_tempering.py:291-398): per chain,
    pairs idx = 0 .. T-2 in order; ratio = (1/T[idx+1] - 1/T[idx]) *
    (L[idx] - L[idx+1]) on the current arrays; accept iff log_u < ratio, then
    the two slots of positions and log-likelihoods are swapped.
    x [C, T, d], loglik [C, T] (untempered), log_u [C, T-1]."""
    x = x.clone()
    L = loglik.clone()
    T = torch.as_tensor(temperatures, dtype=torch.float64)
    n = x.shape[1]
    acc = torch.zeros(x.shape[0], max(n - 1, 0), dtype=torch.bool)
    for idx in range(n - 1):
        ratio = (1.0 / T[idx + 1] - 1.0 / T[idx]) * (L[:, idx] - L[:, idx + 1])
        a = log_u[:, idx] < ratio
        acc[:, idx] = a
        sw = x[a, idx].clone()
        x[a, idx] = x[a, idx + 1]
        x[a, idx + 1] = sw
        sl = L[a, idx].clone()
        L[a, idx] = L[a, idx + 1]
        L[a, idx + 1] = sl
    return x, L, acc


def adapt_temperature(temperatures, do_accept):
    """ParallelTempering._adapt_temperature (parallel_tempering.py:400-436):
    acceptance_rate = mean over chains; damping = (100 / n_chains) *
    (rate[:-1] - rate[1:]); for i = 1 .. T-2,
    new[i] = new[i-1] + (T[i] - T[i-1]) * exp(damping[i-1])."""
    T = torch.as_tensor(temperatures, dtype=torch.float64)
    acc = torch.as_tensor(do_accept, dtype=torch.float64)
    rate = acc.mean(0)
    damping = (100.0 / acc.shape[0]) * (rate[:-1] - rate[1:])
    new = T.clone()
    for i in range(1, T.shape[0] - 1):
        new[i] = new[i - 1] + (T[i] - T[i - 1]) * torch.exp(damping[i - 1])
    return new
