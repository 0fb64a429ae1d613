"""MCMC convergence diagnostics for the position buffers.

Rank-normalised split-R-hat and bulk effective sample size (Vehtari,
Gelman, Simpson, Carpenter, Bürkner, "Rank-normalization, folding, and
localization: an improved R-hat", Bayesian Analysis 2021) — the method of
``az.rhat(..., method="rank")`` used in docs/tutorials/dualmoon.ipynb. The
reference computes no diagnostics itself. Everything runs batched over
dimensions with torch (on the device holding the draws).

Input layout: draws [n_chains, n_draws, n_dim] (the Buffer layout).
"""

from __future__ import annotations

import math

import torch


def _split(x: torch.Tensor) -> torch.Tensor:
    """[C, N, D] -> [2C, N//2, D] (first / second half of every chain)."""
    n = x.shape[1] // 2
    return torch.cat([x[:, :n], x[:, x.shape[1] - n:]], 0)


def _ndtri(p: torch.Tensor) -> torch.Tensor:
    return math.sqrt(2.0) * torch.special.erfinv(2.0 * p - 1.0)


def rank_normalize(x: torch.Tensor) -> torch.Tensor:
    """z = Phi^-1((rank - 3/8) / (S + 1/4)) over all draws of each dim (average ranks for ties)."""
    C, N, D = x.shape
    flat = x.reshape(C * N, D)
    S = C * N
    order = torch.argsort(flat, dim=0, stable=True)
    ranks = torch.empty_like(flat)
    ar = torch.arange(1, S + 1, dtype=flat.dtype, device=flat.device).unsqueeze(1).expand(S, D)
    ranks.scatter_(0, order, ar)
    # average ranks of tied values
    srt = torch.gather(flat, 0, order)
    if bool((srt[1:] == srt[:-1]).any()):
        for j in range(D):
            v, inv, cnt = torch.unique(flat[:, j], return_inverse=True, return_counts=True)
            if int(cnt.max()) > 1:
                sums = torch.zeros(v.shape[0], dtype=flat.dtype, device=flat.device).scatter_add_(0, inv, ranks[:, j])
                ranks[:, j] = (sums / cnt.to(flat.dtype))[inv]
    z = _ndtri((ranks - 0.375) / (S + 0.25))
    return z.reshape(C, N, D)


def _rhat_raw(x: torch.Tensor) -> torch.Tensor:
    """Classic R-hat on [M, N, D] (already split)."""
    M, N, _ = x.shape
    means = x.mean(1)
    var = x.var(1, unbiased=True)
    B = N * means.var(0, unbiased=True)
    W = var.mean(0)
    var_hat = (N - 1) / N * W + B / N
    return torch.sqrt(var_hat / W)


def rhat(x: torch.Tensor) -> torch.Tensor:
    """Rank-normalised split-R-hat per dim: max(bulk, folded-tail)."""
    x = x.to(torch.float64)
    xs = _split(x)
    bulk = _rhat_raw(rank_normalize(xs))
    med = x.reshape(-1, x.shape[-1]).median(0).values
    tail = _rhat_raw(rank_normalize(_split((x - med).abs())))
    return torch.maximum(bulk, tail)


def _autocov(x: torch.Tensor) -> torch.Tensor:
    """Autocovariance along dim 1 via FFT, [M, N, D] -> [M, N, D] (biased, as ArviZ)."""
    M, N, D = x.shape
    xc = x - x.mean(1, keepdim=True)
    nfft = 1 << (2 * N - 1).bit_length()
    f = torch.fft.rfft(xc, n=nfft, dim=1)
    ac = torch.fft.irfft(f * f.conj(), n=nfft, dim=1)[:, :N]
    return ac / N


def ess_raw(x: torch.Tensor) -> torch.Tensor:
    """Effective sample size of [M, N, D] (already split) with Geyer's initial
    monotone sequence on the multi-chain autocorrelation (Vehtari et al. eq. 10)."""
    M, N, D = x.shape
    acov = _autocov(x)
    chain_var = acov[:, 0] * N / (N - 1.0)
    mean_var = chain_var.mean(0)
    var_plus = mean_var * (N - 1.0) / N
    if M > 1:
        var_plus = var_plus + x.mean(1).var(0, unbiased=True)
    rho = 1.0 - (mean_var - acov.mean(0)) / var_plus  # [N, D]
    rho[0] = 1.0
    # Geyer: sums of consecutive pairs P_k = rho_{2k} + rho_{2k+1} while positive, made monotone
    n_pairs = N // 2
    P = rho[: 2 * n_pairs].reshape(n_pairs, 2, D).sum(1)
    positive = P > 0
    # truncate at the first non-positive pair
    first_bad = torch.where(positive.all(0), torch.full((D,), n_pairs, device=x.device),
                            (~positive).to(torch.int64).argmax(0))
    idx = torch.arange(n_pairs, device=x.device).unsqueeze(1)
    keep = idx < first_bad.unsqueeze(0)
    P = torch.where(keep, P, torch.zeros_like(P))
    P = torch.cummin(P, dim=0).values  # monotone (non-increasing) sequence
    P = torch.where(keep, P, torch.zeros_like(P))
    tau = -1.0 + 2.0 * P.sum(0)
    tau = torch.clamp(tau, min=1.0 / math.log10(M * N))
    return M * N / tau


def ess_bulk(x: torch.Tensor) -> torch.Tensor:
    """Bulk-ESS per dim: ESS of the rank-normalised split chains."""
    return ess_raw(rank_normalize(_split(x.to(torch.float64))))


def summary(x: torch.Tensor) -> dict:
    """min bulk-ESS, max R-hat, means and variances over dims of draws [C, N, D]."""
    e = ess_bulk(x)
    r = rhat(x)
    flat = x.reshape(-1, x.shape[-1]).to(torch.float64)
    return {"min_ess_bulk": float(e.min()), "max_rhat": float(r.max()), "mean": flat.mean(0), "var": flat.var(0),
            "ess_bulk": e, "rhat": r}
