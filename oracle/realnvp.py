"""FP64 restatement of RealNVP — src/flowMC/resource/model/nf_model/realNVP.py.

Test infrastructure only (see the package docstring). Follows the reference
literally: per layer a MaskedCouplingLayer (common.py:127-168) around an
MLPAffine (common.py:171-208) whose scale and shift MLPs (common.py:68-112)
see the full vector x * mask and predict all d outputs; the mask keeps the
conditioning dims. The MLPs are [d, hidden, d] with ReLU (the MLP default).

Parameters come as a dict of per-layer tensors {"Ws1","bs1","Ws2","bs2",
"Wt1","bt1","Wt2","bt2"} (eqx.nn.Linear convention y = W x + b), plus
"data_mean","data_cov","base_mean","base_cov"; dt is the MLPAffine factor
(1 in RealNVP).
"""

from __future__ import annotations

import math

import torch

LEAVES = ("Ws1", "bs1", "Ws2", "bs2", "Wt1", "bt1", "Wt2", "bt2")


def mask(d, layer, dtype=torch.float64):
    """realNVP.py:155-158: ones, first int(d/2) zeroed, flipped on even layers
    (1 = conditioning input, 0 = transformed)."""
    m = torch.ones(d, dtype=dtype)
    m[: int(d / 2)] = 0
    return 1 - m if layer % 2 == 0 else m


def mlp(W1, b1, W2, b2, c):
    """common.py:109-112 with layers [Linear, relu, Linear]."""
    return torch.relu(c @ W1.T + b1) @ W2.T + b2


def affine_coupling(lp, x, m, dt=1.0, inverse=False):
    """MaskedCouplingLayer(MLPAffine) forward / inverse, common.py:150-208."""
    c = x * m
    scale = torch.tanh(mlp(lp["Ws1"], lp["bs1"], lp["Ws2"], lp["bs2"], c)) * dt
    shift = mlp(lp["Wt1"], lp["bt1"], lp["Wt2"], lp["bt2"], c) * dt
    if inverse:
        y = x * torch.exp(-scale) - shift
        ld = -scale
    else:
        y = (x + shift) * torch.exp(scale)
        ld = scale
    y = (1 - m) * y + m * x
    return y, ((1 - m) * ld).sum(-1)


def forward(params, x, dt=1.0):
    """RealNVP.forward (realNVP.py:163-180): layers 0 .. L-1."""
    d = x.shape[-1]
    ld = torch.zeros(x.shape[:-1], dtype=x.dtype)
    for l, lp in enumerate(params["layers"]):
        x, li = affine_coupling(lp, x, mask(d, l, x.dtype), dt)
        ld = ld + li
    return x, ld


def inverse(params, y, dt=1.0):
    """RealNVP.inverse (realNVP.py:182-198): layers L-1 .. 0."""
    d = y.shape[-1]
    ld = torch.zeros(y.shape[:-1], dtype=y.dtype)
    for l in reversed(range(len(params["layers"]))):
        y, li = affine_coupling(params["layers"][l], y, mask(d, l, y.dtype), dt, inverse=True)
        ld = ld + li
    return y, ld


def log_prob(params, x, dt=1.0):
    """RealNVP.log_prob (realNVP.py:207-214): standardise with the data
    statistics (diagonal), forward, plus the log-density of N(0, I) — not of
    base_dist, as in the reference."""
    sd = torch.sqrt(torch.diagonal(params["data_cov"]))
    z = (x - params["data_mean"]) / sd
    y, ld = forward(params, z, dt)
    d = x.shape[-1]
    return ld - 0.5 * (y * y).sum(-1) - 0.5 * d * math.log(2 * math.pi)


def sample_from_noise(params, xi, dt=1.0):
    """RealNVP.sample (realNVP.py:200-205) with base_dist = diagonal Gaussian:
    z = base_mean + sqrt(diag base_cov) * xi, inverse, de-standardise."""
    z = params["base_mean"] + torch.sqrt(torch.diagonal(params["base_cov"])) * xi
    x, _ = inverse(params, z, dt)
    return x * torch.sqrt(torch.diagonal(params["data_cov"])) + params["data_mean"]


def loss(params, x, dt=1.0):
    """-mean log_prob (nf_model/base.py:98-100)."""
    return -log_prob(params, x, dt).mean()


def init_params(d, L, hidden, seed=0, dtype=torch.float64, scale=1e-4, out_scale=None):
    """Distributions of the reference initialisation (common.py:84-107): hidden
    Linear weights N(0, scale / d), every bias and the output Linear weights
    U(-1/sqrt(fan_in), 1/sqrt(fan_in)) (eqx.nn.Linear). ``out_scale``
    replaces the output-layer bound (tests use larger values to exercise the
    nonlinearity)."""
    g = torch.Generator().manual_seed(seed)
    layers = []
    for _ in range(L):
        lp = {}
        for p in ("s", "t"):
            lp[f"W{p}1"] = torch.randn(hidden, d, generator=g, dtype=dtype) * math.sqrt(scale / d)
            lp[f"b{p}1"] = (torch.rand(hidden, generator=g, dtype=dtype) * 2 - 1) / math.sqrt(d)
            bo = out_scale if out_scale is not None else 1 / math.sqrt(hidden)
            lp[f"W{p}2"] = (torch.rand(d, hidden, generator=g, dtype=dtype) * 2 - 1) * bo
            lp[f"b{p}2"] = (torch.rand(d, generator=g, dtype=dtype) * 2 - 1) * bo
        layers.append(lp)
    return {"layers": layers, "data_mean": torch.zeros(d, dtype=dtype), "data_cov": torch.eye(d, dtype=dtype),
            "base_mean": torch.zeros(d, dtype=dtype), "base_cov": torch.eye(d, dtype=dtype)}


def to_flat(params) -> torch.Tensor:
    """Flat layout of the device parameter vector: per layer the leaves in
    LEAVES order, then data_mean, data_cov, base_mean, base_cov."""
    parts = []
    for lp in params["layers"]:
        parts += [lp[k].reshape(-1) for k in LEAVES]
    parts += [params[k].reshape(-1) for k in ("data_mean", "data_cov", "base_mean", "base_cov")]
    return torch.cat(parts)


def from_flat(theta, d, L, hidden):
    shapes = {}
    for p in ("s", "t"):
        shapes.update({f"W{p}1": (hidden, d), f"b{p}1": (hidden,), f"W{p}2": (d, hidden), f"b{p}2": (d,)})
    o = 0
    layers = []
    for _ in range(L):
        lp = {}
        for k in LEAVES:
            n = math.prod(shapes[k])
            lp[k] = theta[o:o + n].reshape(shapes[k])
            o += n
        layers.append(lp)
    out = {"layers": layers}
    for k, shp in (("data_mean", (d,)), ("data_cov", (d, d)), ("base_mean", (d,)), ("base_cov", (d, d))):
        n = math.prod(shp)
        out[k] = theta[o:o + n].reshape(shp)
        o += n
    assert o == theta.numel()
    return out
