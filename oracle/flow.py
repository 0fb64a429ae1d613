"""FP64 restatement of the MaskedCouplingRQSpline flow (torch, autograd-able).

Follows the reference literally: the conditioner sees x*mask over all d
inputs and predicts parameters for all d dims, the spline runs on every dim
and the mask then keeps the conditioning dims (common.py:150-168); the bin is
selected with the one-hot formulation of rqSpline.py:63-80.

Parameters come as a dict of per-layer tensors:
  {"scale","shift","W1","b1",...,"W{n+1}","b{n+1}"} per layer for a
  conditioner with n hidden layers (W1 .. Wn hidden, W{n+1} the output
  layer; eqx.nn.Linear convention y = W x + b), plus "data_mean","data_cov",
  "base_mean","base_cov".
"""

from __future__ import annotations

import math

import torch


def normalize_bin_sizes(u, total, min_size):
    """rqSpline.py:20-27."""
    K = u.shape[-1]
    return torch.softmax(u, dim=-1) * (total - K * min_size) + min_size


def normalize_knot_slopes(u, min_slope):
    """rqSpline.py:30-39."""
    off = math.log(math.exp(1.0 - min_slope) - 1.0)
    return torch.nn.functional.softplus(u + off) + min_slope


def spline_params(raw, K, lo=-10.0, hi=10.0, min_bin=1e-4, min_slope=1e-4):
    """RQSpline.get_params (rqSpline.py:310-338): raw [..., 3K+1] -> x_pos, y_pos, slopes [..., K+1]."""
    w = normalize_bin_sizes(raw[..., :K], hi - lo, min_bin)
    h = normalize_bin_sizes(raw[..., K:2 * K], hi - lo, min_bin)
    pad_lo = torch.full(raw.shape[:-1] + (1,), lo, dtype=raw.dtype)
    pad_hi = torch.full(raw.shape[:-1] + (1,), hi, dtype=raw.dtype)
    x_pos = torch.cat([pad_lo, lo + torch.cumsum(w[..., :-1], -1), pad_hi], -1)
    y_pos = torch.cat([pad_lo, lo + torch.cumsum(h[..., :-1], -1), pad_hi], -1)
    s = normalize_knot_slopes(raw[..., 2 * K:], min_slope)
    return x_pos, y_pos, s


def _select_bin(v, pos):
    below = v <= pos[..., 0]
    above = v >= pos[..., -1]
    correct = (v[..., None] >= pos[..., :-1]) & (v[..., None] < pos[..., 1:])
    anyb = correct.any(-1, keepdim=True)
    first = torch.zeros_like(correct)
    first[..., 0] = True
    correct = torch.where(anyb, correct, first).to(v.dtype)
    return below, above, correct


def spline_fwd(x, x_pos, y_pos, d):
    """_rational_quadratic_spline_fwd, rqSpline.py:42-128 (elementwise over leading dims)."""
    below, above, cb = _select_bin(x, x_pos)
    xk = (cb * x_pos[..., :-1]).sum(-1)
    xk1 = (cb * x_pos[..., 1:]).sum(-1)
    yk = (cb * y_pos[..., :-1]).sum(-1)
    yk1 = (cb * y_pos[..., 1:]).sum(-1)
    dk = (cb * d[..., :-1]).sum(-1)
    dk1 = (cb * d[..., 1:]).sum(-1)
    w = xk1 - xk
    hh = yk1 - yk
    s = hh / w
    z = torch.clamp((x - xk) / w, 0.0, 1.0)
    sq_z = z * z
    z1mz = z - sq_z
    sq_1mz = (1.0 - z) ** 2
    st = dk1 + dk - 2.0 * s
    num = hh * (s * sq_z + dk * z1mz)
    den = s + st * z1mz
    y = yk + num / den
    ld = 2.0 * torch.log(s) + torch.log(dk1 * sq_z + 2.0 * s * z1mz + dk * sq_1mz) - 2.0 * torch.log(den)
    y = torch.where(below, (x - x_pos[..., 0]) * d[..., 0] + y_pos[..., 0], y)
    y = torch.where(above, (x - x_pos[..., -1]) * d[..., -1] + y_pos[..., -1], y)
    ld = torch.where(below, torch.log(d[..., 0]), ld)
    ld = torch.where(above, torch.log(d[..., -1]), ld)
    return y, ld


def spline_inv(y, x_pos, y_pos, d):
    """_rational_quadratic_spline_inv + _safe_quadratic_root, rqSpline.py:131-239."""
    below, above, cb = _select_bin(y, y_pos)
    xk = (cb * x_pos[..., :-1]).sum(-1)
    xk1 = (cb * x_pos[..., 1:]).sum(-1)
    yk = (cb * y_pos[..., :-1]).sum(-1)
    yk1 = (cb * y_pos[..., 1:]).sum(-1)
    dk = (cb * d[..., :-1]).sum(-1)
    dk1 = (cb * d[..., 1:]).sum(-1)
    w = xk1 - xk
    hh = yk1 - yk
    s = hh / w
    wv = torch.clamp((y - yk) / hh, 0.0, 1.0)
    st = dk1 + dk - 2.0 * s
    c = -s * wv
    b = dk - st * wv
    a = s - b
    disc = b * b - 4.0 * a * c
    sq = torch.sqrt(torch.clamp(disc, min=torch.finfo(disc.dtype).tiny))
    sq = torch.where(disc > 0, sq, torch.zeros_like(sq))
    z = torch.where(b >= 0, 2.0 * c / (-b - sq), (-b + sq) / (2.0 * a))
    z = torch.clamp(z, 0.0, 1.0)
    x = w * z + xk
    sq_z = z * z
    z1mz = z - sq_z
    sq_1mz = (1.0 - z) ** 2
    den = s + st * z1mz
    ld = -2.0 * torch.log(s) - torch.log(dk1 * sq_z + 2.0 * s * z1mz + dk * sq_1mz) + 2.0 * torch.log(den)
    x = torch.where(below, (y - y_pos[..., 0]) / d[..., 0] + x_pos[..., 0], x)
    x = torch.where(above, (y - y_pos[..., -1]) / d[..., -1] + x_pos[..., -1], x)
    ld = torch.where(below, -torch.log(d[..., 0]), ld)
    ld = torch.where(above, -torch.log(d[..., -1]), ld)
    return x, ld


def mask(d, layer, dtype=torch.float64):
    """rqSpline.py:434: mask[j] = (j + layer) % 2 (1 = conditioning)."""
    return ((torch.arange(d) + layer) % 2).to(dtype)


class _RoundSTE(torch.autograd.Function):
    """Round to BF16 (from the FP32 value, round-to-nearest-even) in forward;
    pass the gradient through unchanged."""

    @staticmethod
    def forward(ctx, x):
        return x.float().bfloat16().to(x.dtype)

    @staticmethod
    def backward(ctx, g):
        return g


class _RoundGrad(torch.autograd.Function):
    """Identity in forward; round the incoming gradient to BF16 in backward."""

    @staticmethod
    def forward(ctx, x):
        return x.clone()

    @staticmethod
    def backward(ctx, g):
        return g.float().bfloat16().to(g.dtype)


_EMULATE_BF16 = [False]


class emulate_bf16:
    """Context manager: evaluate the conditioner the way the BF16 kernels do —
    GEMM operands (input, weights, hidden activations) rounded to BF16 with
    FP32-or-better accumulation and FP32 biases, and the gradients the adjoint
    kernels store in BF16 (w.r.t. spline parameters and hidden
    pre-activations) rounded likewise."""

    def __enter__(self):
        _EMULATE_BF16[0] = True

    def __exit__(self, *a):
        _EMULATE_BF16[0] = False


def n_linear(lp) -> int:
    """Linear layers of a conditioner: its hidden layers plus the output layer."""
    n = 0
    while f"W{n + 1}" in lp:
        n += 1
    return n


def mlp(lp, c):
    """MLP with tanh hidden activations (common.py:109-112, rqSpline.py:428-433):
    any number of hidden layers (common.py:92-107)."""
    n = n_linear(lp)
    if not _EMULATE_BF16[0]:
        h = c
        for i in range(1, n):
            h = torch.tanh(h @ lp[f"W{i}"].T + lp[f"b{i}"])
        return h @ lp[f"W{n}"].T + lp[f"b{n}"]
    r, rg = _RoundSTE.apply, _RoundGrad.apply
    h = c
    for i in range(1, n):
        h = torch.tanh(rg(r(h) @ r(lp[f"W{i}"]).T + lp[f"b{i}"]))
    return rg(r(h) @ r(lp[f"W{n}"]).T + lp[f"b{n}"])


def coupling(lp, x, layer, K, inverse, rng=(-10.0, 10.0)):
    """MaskedCouplingLayer(RQSpline) forward / inverse (common.py:150-168, rqSpline.py:340-361)."""
    d = x.shape[-1]
    m = mask(d, layer, x.dtype)
    raw = mlp(lp, x * m).reshape(x.shape[:-1] + (d, 3 * K + 1))
    x_pos, y_pos, s = spline_params(raw, K, rng[0], rng[1])
    fn = spline_inv if inverse else spline_fwd
    y, ld = fn(x, x_pos, y_pos, s)
    y = (1 - m) * y + m * x
    return y, ((1 - m) * ld).sum(-1)


def affine(lp, x, inverse):
    """ScalarAffine in a MaskedCouplingLayer with an all-zero mask (common.py:211-240)."""
    d = x.shape[-1]
    if inverse:
        return x * torch.exp(-lp["scale"]) - lp["shift"], -lp["scale"] * d * torch.ones(x.shape[:-1], dtype=x.dtype)
    return (x + lp["shift"]) * torch.exp(lp["scale"]), lp["scale"] * d * torch.ones(x.shape[:-1], dtype=x.dtype)


def forward(params, x, K):
    """MaskedCouplingRQSpline.forward (rqSpline.py:450-468)."""
    ld = torch.zeros(x.shape[:-1], dtype=x.dtype)
    for i, lp in enumerate(params["layers"]):
        x, a = affine(lp, x, False)
        x, b = coupling(lp, x, i, K, False)
        ld = ld + a + b
    return x, ld


def inverse(params, y, K, exact=False):
    """MaskedCouplingRQSpline.inverse (rqSpline.py:470-488): layers reversed, but
    inside a layer ScalarAffine^-1 before spline^-1 (the reference's order);
    exact=True applies spline^-1 first."""
    ld = torch.zeros(y.shape[:-1], dtype=y.dtype)
    n = len(params["layers"])
    for i in reversed(range(n)):
        lp = params["layers"][i]
        if exact:
            y, b = coupling(lp, y, i, K, True)
            y, a = affine(lp, y, True)
        else:
            y, a = affine(lp, y, True)
            y, b = coupling(lp, y, i, K, True)
        ld = ld + a + b
    return y, ld


def base_log_prob(params, y):
    """Gaussian.log_prob = multivariate_normal.logpdf(y, mean, cov) (common.py:285-286)."""
    mvn = torch.distributions.MultivariateNormal(params["base_mean"], covariance_matrix=params["base_cov"])
    return mvn.log_prob(y)


def log_prob(params, x, K, learnable_base=False):
    """MaskedCouplingRQSpline.log_prob (rqSpline.py:498-504): no -sum log sigma term.
    data_mean / data_cov are stop_gradient leaves (nf_model/base.py:25-31); so
    are the base mean / cov unless the base is learnable (common.py:261-273).
    A learnable covariance enters through jax.scipy.stats.multivariate_normal,
    whose Cholesky factorisation (lax.linalg.cholesky) symmetrises its input:
    the density sees (cov + cov^T) / 2."""
    std = torch.sqrt(torch.diagonal(params["data_cov"].detach()))
    z = (x - params["data_mean"].detach()) / std
    y, ld = forward(params, z, K)
    if learnable_base:
        c = params["base_cov"]
        base = {"base_mean": params["base_mean"], "base_cov": 0.5 * (c + c.transpose(-1, -2))}
    else:
        base = {"base_mean": params["base_mean"].detach(), "base_cov": params["base_cov"].detach()}
    return ld + base_log_prob(base, y)


def sample_from_noise(params, xi, K, exact=False):
    """MaskedCouplingRQSpline.sample (rqSpline.py:490-496) with injected base noise xi ~ N(0, I):
    base draw = base_mean + chol(base_cov) xi (jax.random.multivariate_normal)."""
    L = torch.linalg.cholesky(params["base_cov"])
    z = params["base_mean"] + xi @ L.T
    x, _ = inverse(params, z, K, exact)
    return x * torch.sqrt(torch.diagonal(params["data_cov"])) + params["data_mean"]


def loss(params, x, K):
    """NFModel.loss_fn (nf_model/base.py:98-100)."""
    return -log_prob(params, x, K).mean()


def _hidden(h1, h2):
    """Hidden widths from (h1, h2) or from a list passed as h1 (h2 None)."""
    if h2 is None:
        return [int(h) for h in h1]
    return [int(h1), int(h2)]


def _shapes(d, K, hidden):
    P = 3 * K + 1
    ins = [d] + hidden
    outs = hidden + [d * P]
    shp = [("scale", ()), ("shift", ())]
    for i, (a, b) in enumerate(zip(ins, outs)):
        shp += [(f"W{i + 1}", (b, a)), (f"b{i + 1}", (b,))]
    return shp


def init_params(d, L, K, h1, h2=None, seed=0, dtype=torch.float64, scale=0.0, hidden_var=1.0, out_var=1.0):
    """Random parameters for tests (not the reference's init draws). Hidden
    widths are (h1, h2) or the list h1. ``scale`` sets the ScalarAffine
    scale/shift magnitude (0 = the reference's init); hidden weights are
    N(0, hidden_var / fan_in) (the reference uses 1e-2, rqSpline.py:431 with
    common.py:98-100); output weights N(0, out_var / fan_in)."""
    hidden = _hidden(h1, h2)
    g = torch.Generator().manual_seed(seed)
    n = len(hidden) + 1
    layers = []
    for _ in range(L):
        lp = {"scale": (torch.rand((), generator=g, dtype=dtype) - 0.5) * 2 * scale,
              "shift": (torch.rand((), generator=g, dtype=dtype) - 0.5) * 2 * scale}
        for k, shp in _shapes(d, K, hidden)[2:]:
            i = int(k[1:])
            if k[0] == "W":
                var = hidden_var if i < n else out_var
                lp[k] = torch.randn(*shp, generator=g, dtype=dtype) * math.sqrt(var / shp[1])
            else:
                lp[k] = (torch.rand(*shp, generator=g, dtype=dtype) - 0.5) * (0.2 if i < n else 0.5)
        layers.append(lp)
    return {"layers": layers, "data_mean": torch.zeros(d, dtype=dtype), "data_cov": torch.eye(d, dtype=dtype),
            "base_mean": torch.zeros(d, dtype=dtype), "base_cov": torch.eye(d, dtype=dtype)}


def to_flat(params) -> torch.Tensor:
    """Flatten in the master-parameter order of the device flow (the leaf
    order of the reference's pytree: scale, shift, then weight and bias of
    every linear layer)."""
    parts = []
    for lp in params["layers"]:
        parts += [lp["scale"].reshape(-1), lp["shift"].reshape(-1)]
        for i in range(1, n_linear(lp) + 1):
            parts += [lp[f"W{i}"].reshape(-1), lp[f"b{i}"].reshape(-1)]
    parts += [params["data_mean"].reshape(-1), params["data_cov"].reshape(-1), params["base_mean"].reshape(-1),
              params["base_cov"].reshape(-1)]
    return torch.cat(parts)


def from_flat(theta, d, L, K, h1, h2=None):
    shapes = _shapes(d, K, _hidden(h1, h2))
    o = 0
    layers = []
    for _ in range(L):
        lp = {}
        for k, shp in shapes:
            n = math.prod(shp) if shp else 1
            lp[k] = theta[o:o + n].reshape(shp)
            o += n
        layers.append(lp)
    out = {"layers": layers}
    for k, n, shp in (("data_mean", d, (d,)), ("data_cov", d * d, (d, d)), ("base_mean", d, (d,)),
                      ("base_cov", d * d, (d, d))):
        out[k] = theta[o:o + n].reshape(shp)
        o += n
    assert o == theta.numel()
    return out
