"""Plain FP32 per-step GPU path — the comparison baseline of the NF-MCMC benchmark.

The flowMC algorithm written in eager PyTorch on the GPU, one library op per
launch, FP32 throughout, keeping the reference's per-step work:
  * MALA re-evaluates value-and-gradient at the current point every step
    (src/flowMC/resource/kernel/MALA.py:59,71-73) — 2 evaluations per step;
  * the conditioner runs at full width (all d inputs, d*(3K+1) outputs) and
    the spline selects bins with the one-hot formulation
    (rqSpline.py:63-80, common.py:150-168);
  * training uses torch autograd + clip_grad_norm_(1.0) + torch.optim.AdamW
    (the optax chain of src/flowMC/resource/optimizer.py:19-22);
  * the NF independence-MH scan is a Python loop over proposals
    (NF_proposal.py:91-126).
It is used only by bench.py; the product path is flowmc_amd.
"""

from __future__ import annotations

import sys
import math

import torch


class PlainFlow:
    def __init__(self, theta: torch.Tensor, d, L, K, h1, h2, device="cuda"):
        self.d, self.L, self.K, self.h1, self.h2 = d, L, K, h1, h2
        P = 3 * K + 1
        shapes = [("scale", ()), ("shift", ()), ("W1", (h1, d)), ("b1", (h1,)), ("W2", (h2, h1)), ("b2", (h2,)),
                  ("W3", (d * P, h2)), ("b3", (d * P,))]
        o = 0
        self.layers = []
        th = theta.detach().to(device, torch.float32)
        for _ in range(L):
            lp = {}
            for k, shp in shapes:
                n = math.prod(shp) if shp else 1
                lp[k] = th[o:o + n].reshape(shp).clone().requires_grad_(True)
                o += n
            self.layers.append(lp)
        self.data_mean = th[o:o + d].clone()
        self.data_cov = th[o + d:o + d + d * d].reshape(d, d).clone()
        o += d + d * d
        self.base_mean = th[o:o + d].clone()
        self.base_cov = th[o + d:o + d + d * d].reshape(d, d).clone()
        self.masks = [((torch.arange(d, device=device) + i) % 2).float() for i in range(L)]

    def params(self):
        return [t for lp in self.layers for t in lp.values()]

    def _spline_params(self, raw):
        K = self.K
        lo, hi = -10.0, 10.0
        w = torch.softmax(raw[..., :K], -1) * (20.0 - K * 1e-4) + 1e-4
        h = torch.softmax(raw[..., K:2 * K], -1) * (20.0 - K * 1e-4) + 1e-4
        pad_lo = torch.full(raw.shape[:-1] + (1,), lo, device=raw.device)
        pad_hi = torch.full(raw.shape[:-1] + (1,), hi, device=raw.device)
        xp = torch.cat([pad_lo, lo + torch.cumsum(w[..., :-1], -1), pad_hi], -1)
        yp = torch.cat([pad_lo, lo + torch.cumsum(h[..., :-1], -1), pad_hi], -1)
        s = torch.nn.functional.softplus(raw[..., 2 * K:] + math.log(math.exp(1 - 1e-4) - 1)) + 1e-4
        return xp, yp, s

    @staticmethod
    def _bin(v, pos):
        below = v <= pos[..., 0]
        above = v >= pos[..., -1]
        cb = ((v[..., None] >= pos[..., :-1]) & (v[..., None] < pos[..., 1:]))
        anyb = cb.any(-1, keepdim=True)
        first = torch.zeros_like(cb)
        first[..., 0] = True
        return below, above, torch.where(anyb, cb, first).float()

    def _spline(self, x, raw, inverse):
        xp, yp, d = self._spline_params(raw)
        below, above, cb = self._bin(x, yp if inverse else xp)
        xk, xk1 = (cb * xp[..., :-1]).sum(-1), (cb * xp[..., 1:]).sum(-1)
        yk, yk1 = (cb * yp[..., :-1]).sum(-1), (cb * yp[..., 1:]).sum(-1)
        dk, dk1 = (cb * d[..., :-1]).sum(-1), (cb * d[..., 1:]).sum(-1)
        w, hh = xk1 - xk, yk1 - yk
        s = hh / w
        st = dk1 + dk - 2 * s
        if not inverse:
            z = torch.clamp((x - xk) / w, 0, 1)
        else:
            wv = torch.clamp((x - yk) / hh, 0, 1)
            c = -s * wv
            b = dk - st * wv
            a = s - b
            disc = b * b - 4 * a * c
            sq = torch.where(disc > 0, torch.sqrt(torch.clamp(disc, min=1.17549435e-38)), torch.zeros_like(disc))
            z = torch.clamp(torch.where(b >= 0, 2 * c / (-b - sq), (-b + sq) / (2 * a)), 0, 1)
        zz = z - z * z
        den = s + st * zz
        ld = 2 * torch.log(s) + torch.log(dk1 * z * z + 2 * s * zz + dk * (1 - z) ** 2) - 2 * torch.log(den)
        if not inverse:
            y = yk + hh * (s * z * z + dk * zz) / den
            y = torch.where(below, (x - xp[..., 0]) * d[..., 0] + yp[..., 0], y)
            y = torch.where(above, (x - xp[..., -1]) * d[..., -1] + yp[..., -1], y)
            ld = torch.where(below, torch.log(d[..., 0]), torch.where(above, torch.log(d[..., -1]), ld))
        else:
            y = w * z + xk
            y = torch.where(below, (x - yp[..., 0]) / d[..., 0] + xp[..., 0], y)
            y = torch.where(above, (x - yp[..., -1]) / d[..., -1] + xp[..., -1], y)
            ld = torch.where(below, torch.log(d[..., 0]), torch.where(above, torch.log(d[..., -1]), ld))
            ld = -ld
        return y, ld

    def _coupling(self, i, x, inverse):
        lp, m = self.layers[i], self.masks[i]
        h = torch.tanh(torch.nn.functional.linear(x * m, lp["W1"], lp["b1"]))
        h = torch.tanh(torch.nn.functional.linear(h, lp["W2"], lp["b2"]))
        raw = torch.nn.functional.linear(h, lp["W3"], lp["b3"]).reshape(x.shape[0], self.d, 3 * self.K + 1)
        y, ld = self._spline(x, raw, inverse)
        return (1 - m) * y + m * x, ((1 - m) * ld).sum(-1)

    def forward(self, x):
        ld = torch.zeros(x.shape[0], device=x.device)
        for i, lp in enumerate(self.layers):
            x = (x + lp["shift"]) * torch.exp(lp["scale"])
            ld = ld + lp["scale"] * self.d
            x, l2 = self._coupling(i, x, False)
            ld = ld + l2
        return x, ld

    def inverse(self, y, exact=False):
        """Layers in reverse; inside a layer ScalarAffine^-1 then spline^-1 (the
        reference's order, rqSpline.py:479-485) or, with exact, spline^-1 first."""
        ld = torch.zeros(y.shape[0], device=y.device)
        for i in reversed(range(self.L)):
            lp = self.layers[i]
            if exact:
                y, l2 = self._coupling(i, y, True)
                ld = ld + l2
            y = y * torch.exp(-lp["scale"]) - lp["shift"]
            ld = ld - lp["scale"] * self.d
            if not exact:
                y, l2 = self._coupling(i, y, True)
                ld = ld + l2
        return y, ld

    def log_prob(self, x):
        std = torch.sqrt(torch.diagonal(self.data_cov))
        y, ld = self.forward((x - self.data_mean) / std)
        var = torch.diagonal(self.base_cov)
        u = y - self.base_mean
        return ld - 0.5 * (u * u / var).sum(-1) - 0.5 * torch.log(var).sum() - 0.5 * self.d * math.log(2 * math.pi)

    def sample(self, n, gen, exact=False):
        z = self.base_mean + torch.randn(n, self.d, device=self.base_mean.device, generator=gen) * torch.sqrt(
            torch.diagonal(self.base_cov))
        x, _ = self.inverse(z, exact)
        return x * torch.sqrt(torch.diagonal(self.data_cov)) + self.data_mean


def gmm_logpdf_factory(means, sigmas):
    """Per-point 8-mode diagonal Gaussian mixture log-density written in torch ops."""
    K = means.shape[0]
    logw = torch.full((K,), -math.log(K), device=means.device)
    c = logw - torch.log(sigmas).sum(-1)

    def f(x):
        z = (x - means) / sigmas
        return torch.logsumexp(c - 0.5 * (z * z).sum(-1), 0)

    return f


def run_plain_nfmcmc(theta, cfg, target, x0, seed=0, time_budget_s=None):
    """The RQSpline_MALA bundle's strategy order (RQSpline_MALA.py:251-274) in eager torch.
    Returns (production positions [C, T, d], number of rounds run)."""
    dev = x0.device
    C, d = x0.shape
    g = torch.Generator(device=dev).manual_seed(seed)
    flow = PlainFlow(theta, d, cfg["n_layers"], cfg["n_bins"], cfg["hidden"][0], cfg["hidden"][1], dev)
    opt = torch.optim.AdamW(flow.params(), lr=cfg["lr"], betas=(0.9, 0.999), eps=1e-8, weight_decay=1e-4)
    f = gmm_logpdf_factory(target.means.to(dev, torch.float32), target.sigmas.to(dev, torch.float32))
    vg = torch.func.vmap(torch.func.grad_and_value(f))
    n_local, n_global = cfg["n_local"], cfg["n_global"]
    n_rounds = cfg["n_loops"]
    T = (n_local + n_global) * n_rounds
    train_pos = torch.full((C, T, d), -float("inf"), device=dev)
    prod_pos = torch.full((C, T, d), -float("inf"), device=dev)
    x = x0.clone()
    dt = cfg["step_size"]

    def local(x, buf, cur):
        for s in range(n_local):
            g0, lp0 = vg(x)  # value_and_grad at the current point (MALA.py:59)
            prop = x + dt * dt * g0 / 2 + dt * torch.randn(x.shape, device=dev, generator=g)
            g1, lp1 = vg(prop)
            ratio = lp1 - lp0 + ((prop - x - dt * dt / 2 * g0) ** 2).sum(-1) / (2 * dt * dt) - (
                (x - prop - dt * dt / 2 * g1) ** 2).sum(-1) / (2 * dt * dt)
            acc = torch.log(torch.rand(C, device=dev, generator=g)) < ratio
            x = torch.where(acc[:, None], prop, x)
            buf[:, cur + s] = x
        return x

    def train(buf, filled):
        data = buf[:, :filled][:, -cfg.get("history_window", 100):].reshape(-1, d)  # train_model.py:66-80
        data = data[torch.randint(0, data.shape[0], (cfg["n_max_examples"],), device=dev, generator=g)]
        flow.data_mean = data.mean(0)
        flow.data_cov = torch.cov(data.T)
        best, best_loss = None, 1e9
        for _ in range(cfg["n_epochs"]):
            perm = torch.randperm(data.shape[0], device=dev, generator=g)
            nb = data.shape[0] // cfg["batch_size"]
            for b in range(max(nb, 1)):
                batch = data[perm[b * cfg["batch_size"]:(b + 1) * cfg["batch_size"]]] if nb else data
                loss = -flow.log_prob(batch).mean()
                opt.zero_grad()
                loss.backward()
                torch.nn.utils.clip_grad_norm_(flow.params(), 1.0)
                opt.step()
            v = float(loss)
            if v < best_loss:
                best_loss, best = v, [t.detach().clone() for t in flow.params()]
        if best is not None:
            with torch.no_grad():
                for t, b in zip(flow.params(), best):
                    t.copy_(b)

    def global_step(x, buf, cur):
        with torch.no_grad():
            prop = flow.sample(C * n_global, g, cfg.get("exact_inverse", False)).reshape(C, n_global, d)
            qp = flow.log_prob(prop.reshape(-1, d)).reshape(C, n_global)
            qc = flow.log_prob(x)
            lpp = torch.func.vmap(f)(prop.reshape(-1, d)).reshape(C, n_global)
            lpc = torch.func.vmap(f)(x)
            for s in range(n_global):
                ratio = (lpp[:, s] - lpc) - (qp[:, s] - qc)
                acc = torch.log(torch.rand(C, device=dev, generator=g)) < ratio
                x = torch.where(acc[:, None], prop[:, s], x)
                lpc = torch.where(acc, lpp[:, s], lpc)
                qc = torch.where(acc, qp[:, s], qc)
                buf[:, cur + s] = x
        return x

    import time

    t0 = time.perf_counter()
    cur = 0
    rounds = 0
    for _ in range(n_rounds):
        x = local(x, train_pos, cur)
        cur += n_local
        train(train_pos, cur)
        x = global_step(x, train_pos, cur)
        cur += n_global
        rounds += 1
        print(f"plain path: training round {rounds}/{n_rounds} done at {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
        if time_budget_s is not None and time.perf_counter() - t0 > time_budget_s:
            break
    cur = 0
    prod_rounds = 0
    for _ in range(n_rounds):
        x = local(x, prod_pos, cur)
        cur += n_local
        x = global_step(x, prod_pos, cur)
        cur += n_global
        prod_rounds += 1
        print(f"plain path: production round {prod_rounds}/{n_rounds} done at {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
        if time_budget_s is not None and time.perf_counter() - t0 > 2 * time_budget_s:
            break
    return prod_pos[:, :cur], rounds, prod_rounds
