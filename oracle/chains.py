"""FP64 restatement of the local kernels, their step driver and the NF MH scan.

All randomness is injected: noise [S, C, d] and log-uniforms [S, C], the
same arrays the HIP kernels read in injected-noise mode.
"""

from __future__ import annotations

import math

import torch

from .targets import value_and_grad


def mala_step(x, fn, dt, xi, log_u):
    """One MALA step for a batch of chains — src/flowMC/resource/kernel/MALA.py:26-89.

    As in the reference, (logp, grad) of the current point are re-evaluated
    (MALA.py:59, first scan iteration) rather than taken from the carry, and
    the proposal densities are isotropic Gaussians of variance dt^2 whose
    normalising constants cancel (MALA.py:75-81).
    """
    dt2 = dt * dt
    lp0, g0 = value_and_grad(fn, x)
    prop = x + dt2 * g0 / 2 + dt * xi
    lp1, g1 = value_and_grad(fn, prop)
    d = x.shape[-1]

    def log_n(a, mean):
        y = a - mean
        return -0.5 * (y * y).sum(-1) / dt2 - d / 2 * (math.log(2 * math.pi) + math.log(dt2))

    ratio = lp1 - lp0
    ratio = ratio - log_n(prop, x + dt2 * g0 / 2)
    ratio = ratio + log_n(x, prop + dt2 * g1 / 2)
    acc = log_u < ratio
    x_new = torch.where(acc[..., None], prop, x)
    lp_new = torch.where(acc, lp1, lp0)
    return x_new, lp_new, acc, ratio


def hmc_step(x, lp, fn, eps, n_leapfrog, metric, xi, log_u):
    """One HMC step — src/flowMC/resource/kernel/HMC.py:98-151.

    metric is the condition matrix M: a vector (its diagonal; ones for eye)
    or a dense [d, d] matrix. Momentum = xi @ chol(inv(M)).T (HMC.py:133-136),
    i.e. xi / sqrt(m) for a diagonal M; kinetic = 1/2 sum(p^2 * M) with NumPy
    broadcasting (HMC.py:118-121), which for a dense M is
    1/2 sum_j p_j^2 * colsum_j(M), so its p-gradient (the drift direction,
    HMC.py:74-76) is p * colsum(M); leapfrog coefficient table HMC.py:47-50
    scanned n_leapfrog+2 times (HMC.py:91-95); H0 uses the passed-in
    log_prob (HMC.py:137).
    """
    coefs = torch.ones(n_leapfrog + 2, 2, dtype=x.dtype)
    coefs[0] = torch.tensor([0.0, 0.5])
    coefs[-1] = torch.tensor([1.0, 0.5])
    metric = torch.as_tensor(metric, dtype=x.dtype)
    if metric.ndim == 2:
        chol = torch.linalg.cholesky(torch.linalg.inv(metric))
        p = xi @ chol.T
        kin_w = metric.sum(0)  # (p**2 * M).sum() = sum_j p_j^2 sum_i M_ij
    else:
        p = xi / torch.sqrt(metric)
        kin_w = metric

    def kinetic(p):
        return 0.5 * (p * p * kin_w).sum(-1)

    H = -lp + kinetic(p)
    xl, pl = x, p
    for i in range(n_leapfrog + 2):
        xl = xl + eps * coefs[i, 0] * (pl * kin_w)
        _, g = value_and_grad(fn, xl)
        pl = pl - eps * coefs[i, 1] * (-g)
    lp_prop, _ = value_and_grad(fn, xl)
    proposed_ham = -lp_prop + kinetic(pl)
    log_acc = H - proposed_ham
    acc = log_u < log_acc
    x_new = torch.where(acc[..., None], xl, x)
    lp_new = torch.where(acc, lp_prop, lp)
    return x_new, lp_new, acc, log_acc


def grw_step(x, lp, fn, step_size, xi, log_u):
    """One Gaussian random walk step — src/flowMC/resource/kernel/Gaussian_random_walk.py:25-61:
    proposal = x + step_size * xi; accept iff log u < logp(proposal) - log_prob,
    with the passed-in log_prob of the current point."""
    prop = x + step_size * xi
    lp1, _ = value_and_grad(fn, prop)
    ratio = lp1 - lp
    acc = log_u < ratio
    x_new = torch.where(acc[..., None], prop, x)
    lp_new = torch.where(acc, lp1, lp)
    return x_new, lp_new, acc, ratio


def take_serial_steps(kind, x0, fn, noise, log_u, thinning=1, **kw):
    """TakeSerialSteps.sample + TakeSteps thinning — src/flowMC/strategy/take_steps.py:134-180.

    Returns the thinned (positions [C, R, d], log_probs [C, R], accepts [C, R]),
    the position the strategy hands to the next one (its last thinned row) and
    the log acceptance ratios [C, S] of every step (for margin checks).
    """
    S = noise.shape[0]
    x = x0.clone()
    lp, _ = value_and_grad(fn, x)  # take_steps.py:177
    pos, lps, accs, ratios = [], [], [], []
    for s in range(S):
        if kind == "mala":
            x, lp, acc, r = mala_step(x, fn, kw["step_size"], noise[s], log_u[s])
        elif kind == "hmc":
            x, lp, acc, r = hmc_step(x, lp, fn, kw["step_size"], kw["n_leapfrog"], kw["metric"], noise[s],
                                     log_u[s])
        elif kind == "grw":
            x, lp, acc, r = grw_step(x, lp, fn, kw["step_size"], noise[s], log_u[s])
        else:
            raise ValueError(kind)
        pos.append(x)
        lps.append(lp)
        accs.append(acc)
        ratios.append(r)
    pos = torch.stack(pos, 1)[:, ::thinning]
    lps = torch.stack(lps, 1)[:, ::thinning]
    accs = torch.stack(accs, 1)[:, ::thinning].to(x.dtype)
    return pos, lps, accs, pos[:, -1], torch.stack(ratios, 1)


def nf_mh_scan(x0, lp0, q0, prop, lp_prop, q_prop, log_u):
    """Independence MH over per-chain flow proposals — src/flowMC/resource/kernel/NF_proposal.py:91-126.

    prop [C, S, d], lp_prop / q_prop / log_u [C, S]. Returns positions [C, S, d],
    log_probs [C, S], accepts [C, S] (the scan outputs).
    """
    C, S, _ = prop.shape
    x, lp, q = x0.clone(), lp0.clone(), q0.clone()
    pos, lps, accs = [], [], []
    for s in range(S):
        ratio = (lp_prop[:, s] - lp) - (q_prop[:, s] - q)
        acc = log_u[:, s] < ratio
        x = torch.where(acc[:, None], prop[:, s], x)
        lp = torch.where(acc, lp_prop[:, s], lp)
        q = torch.where(acc, q_prop[:, s], q)
        pos.append(x)
        lps.append(lp)
        accs.append(acc)
    return torch.stack(pos, 1), torch.stack(lps, 1), torch.stack(accs, 1)


def dynamic_update_slice(buf, update, start, axis):
    """jax.lax.dynamic_update_slice_in_dim semantics used by Buffer.update_buffer
    (src/flowMC/resource/buffers.py:32-41): the start index is clamped so the
    update fits inside the buffer."""
    n = buf.shape[axis]
    k = update.shape[axis]
    if k > n:
        raise ValueError("update larger than buffer")
    start = max(0, min(int(start), n - k))
    out = buf.clone()
    idx = [slice(None)] * buf.ndim
    idx[axis] = slice(start, start + k)
    out[tuple(idx)] = update
    return out
