"""RealNVP and AffineCoupling — src/flowMC/resource/model/nf_model/realNVP.py:18-228.

All arithmetic runs in the RealNVP kernels of libflowmc_hip (fmc_rnvp_pass,
fmc_rnvp_backward_layer, fmc_rnvp_weight_grad, fmc_adamw_clip). Parameters
live in one flat FP32 vector ``theta``: per layer the scale MLP (W1 [h, d],
b1 [h], W2 [d, h], b2 [d]) and then the shift MLP, followed by data_mean,
data_cov, base_mean and base_cov. The coupling masks of realNVP.py:155-158
are kept as conditioning / transformed index lists on the device.
"""

from __future__ import annotations

import ctypes as C
import json
import math
from typing import Optional

import numpy as np
import torch

from .... import _native
from .... import random as frandom
from ..common import Gaussian
from .base import NFModel

FORWARD, INVERSE = 0, 1
STANDARDIZE, BASE_LOGP, SAMPLE_BASE = 1, 2, 4


def realnvp_mask(n_features: int, layer: int) -> torch.Tensor:
    """realNVP.py:155-158: 1 = conditioning input, 0 = transformed."""
    m = torch.ones(n_features)
    m[: int(n_features / 2)] = 0
    return 1 - m if layer % 2 == 0 else m


def _layout(masks, d, device) -> torch.Tensor:
    rows = []
    for m in masks:
        cond = [j for j in range(d) if m[j] > 0.5]
        trans = [j for j in range(d) if m[j] <= 0.5]
        rows.append([len(cond), len(trans)] + cond + [0] * (d - len(cond)) + trans + [0] * (d - len(trans)))
    return torch.tensor(rows, dtype=torch.int32, device=device).reshape(-1).contiguous()


class _Workspace:
    """Device scratch of one training batch size (allocated once, reused)."""

    def __init__(self, model: "_RnvpBase", B: int):
        dev = model.theta.device
        d, L = model.n_features, model.n_layers
        self.saved = torch.empty((L, B, d), dtype=torch.float32, device=dev)
        self.logp = torch.empty(B, dtype=torch.float32, device=dev)
        self.g_logp = torch.full((B,), -1.0 / B, dtype=torch.float32, device=dev)
        self.gy = torch.empty((B, d), dtype=torch.float32, device=dev)
        self.gx = torch.empty((B, d), dtype=torch.float32, device=dev)
        self.n_tiles = (B + model._tile - 1) // model._tile
        # tile rows plus the reduction's chunk rows (at most 16)
        self.part = torch.empty((self.n_tiles + 16, model._part_per_tile), dtype=torch.float32, device=dev)
        self.grad = torch.zeros(model.theta.numel(), dtype=torch.float32, device=dev)
        self.norm_ws = torch.empty(1024, dtype=torch.float64, device=dev)
        self.gnorm = torch.empty(1, dtype=torch.float32, device=dev)


class _RnvpBase(NFModel):
    """Device plumbing shared by RealNVP and the stand-alone AffineCoupling."""

    def _configure(self):
        self.layout = _layout(self._masks, self._n_features, self.device)
        probe = self._struct(theta_ptr=1)
        th, pk, pp, tile = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32()
        _native.call("fmc_rnvp_sizes", C.byref(probe), C.byref(th), C.byref(pk), C.byref(pp), C.byref(tile))
        self._theta_size, self._part_per_tile, self._tile = th.value, pp.value, tile.value
        self.packed = torch.empty(pk.value, dtype=torch.float32, device=self.device)
        d, h = self._n_features, self.n_hidden
        self._mlp_size = 2 * h * d + h + d
        self._off = self.n_layers * 2 * self._mlp_size

    def _struct(self, theta_ptr=None) -> _native.fmc_rnvp:
        pk = getattr(self, "packed", None)
        tp = theta_ptr if theta_ptr is not None else _native.dptr(self.theta, name="theta")
        return _native.fmc_rnvp(self._n_features, self.n_layers, self.n_hidden, float(self.dt), tp,
                                pk.data_ptr() if (pk is not None and theta_ptr is None) else None,
                                self.layout.data_ptr())

    def _init_theta(self, key, scale: float, kwargs) -> torch.Tensor:
        """Distributions of common.py:84-107: hidden Linear weights
        N(0, scale / fan_in), its bias and the output Linear
        U(-1/sqrt(fan_in), 1/sqrt(fan_in)) (eqx.nn.Linear). Draws come from a
        torch generator seeded from ``key`` (not JAX's streams)."""
        d, h = self._n_features, self.n_hidden
        k0, k1 = frandom.key_words(key)
        g = torch.Generator().manual_seed((k0 << 32) | k1)
        parts = []
        for _ in range(self.n_layers):
            for _p in range(2):
                parts += [torch.randn(h, d, generator=g) * math.sqrt(scale / d),
                          (torch.rand(h, generator=g) * 2 - 1) / math.sqrt(d),
                          (torch.rand(d, h, generator=g) * 2 - 1) / math.sqrt(h),
                          (torch.rand(d, generator=g) * 2 - 1) / math.sqrt(h)]
        dm = kwargs.get("data_mean")
        dc = kwargs.get("data_cov")
        bd = kwargs.get("base_dist")
        dm = torch.zeros(d) if dm is None else torch.as_tensor(dm, dtype=torch.float32).reshape(d).cpu()
        dc = torch.eye(d) if dc is None else torch.as_tensor(dc, dtype=torch.float32).reshape(d, d).cpu()
        if bd is None:
            bm, bc = torch.zeros(d), torch.eye(d)
        else:
            bm = torch.as_tensor(bd.mean, dtype=torch.float32).reshape(d).cpu()
            bc = torch.as_tensor(bd.cov, dtype=torch.float32).reshape(d, d).cpu()
            if torch.count_nonzero(bc - torch.diag(torch.diagonal(bc))) != 0:
                raise NotImplementedError("the RealNVP kernels need a diagonal base covariance")
        parts += [dm.reshape(-1), dc.reshape(-1), bm.reshape(-1), bc.reshape(-1)]
        return torch.cat([p.reshape(-1).to(torch.float32) for p in parts])

    # ------------------------------------------------------------- views --
    def view(self, layer: int, name: str) -> torch.Tensor:
        """One leaf of ``layer`` inside theta: {W,b}{s,t}{1,2} (s = scale MLP, t = shift MLP)."""
        d, h = self._n_features, self.n_hidden
        shapes = {"W1": (h, d), "b1": (h,), "W2": (d, h), "b2": (d,)}
        p = {"s": 0, "t": 1}[name[1]]
        o = layer * 2 * self._mlp_size + p * self._mlp_size
        for k, shp in shapes.items():
            n = int(np.prod(shp))
            if k == name[0] + name[2]:
                return self.theta[o:o + n].view(shp)
            o += n
        raise KeyError(name)

    def _stat(self, name):
        d = self._n_features
        o = self._off + {"data_mean": 0, "data_cov": d, "base_mean": d + d * d, "base_cov": 2 * d + d * d}[name]
        n = d if name.endswith("mean") else d * d
        t = self.theta[o:o + n]
        return t.view(d) if name.endswith("mean") else t.view(d, d)

    @property
    def data_mean(self):
        return self._stat("data_mean")

    @property
    def data_cov(self):
        return self._stat("data_cov")

    @property
    def base_dist(self):
        return Gaussian(self._stat("base_mean"), self._stat("base_cov"), learnable=False)

    def set_data_stats(self, mean, cov):
        self.data_mean.copy_(mean)
        self.data_cov.copy_(cov)
        self._dirty = True

    def mark_updated(self):
        """Call after editing theta (or a view of it) in place."""
        self._dirty = True

    def clone(self, theta: Optional[torch.Tensor] = None):
        m = type(self).__new__(type(self))
        m.__dict__.update({k: v for k, v in self.__dict__.items() if k not in ("theta", "packed", "_ws")})
        m.packed = torch.empty_like(self.packed)
        m.theta = (self.theta if theta is None else theta).clone()
        m._dirty = True
        m._ws = {}
        return m

    # ----------------------------------------------------------- kernels --
    def _ensure_packed(self):
        if self._dirty:
            s = self._struct()
            _native.call("fmc_rnvp_pack", _native.stream_ptr(), C.byref(s))
            self._dirty = False

    def _pass(self, direction, flags, x, n, key=(0, 0), saved=None, g_logp=None, gy_out=None, want_x=True):
        self._ensure_packed()
        d = self._n_features
        out = torch.empty((n, d), dtype=torch.float32, device=self.device) if want_x else None
        ld = torch.empty(n, dtype=torch.float32, device=self.device)
        s = self._struct()
        _native.call("fmc_rnvp_pass", _native.stream_ptr(), C.byref(s), direction, flags,
                     _native.dptr(x, name="x"), n, key[0], key[1], 0, _native.dptr(out, name="x_out"),
                     _native.dptr(ld, name="logdet"), _native.dptr(saved, name="saved"),
                     _native.dptr(g_logp, name="g_logp"), _native.dptr(gy_out, name="gy_out"))
        return out, ld

    def _as_batch(self, x):
        x = torch.as_tensor(x, dtype=torch.float32)
        if not x.is_cuda:
            x = x.to(self.device)
        lead = x.shape[:-1]
        return x.reshape(-1, self._n_features).contiguous(), lead

    def forward(self, x, key=None, condition=None):
        """(y, log_det), no standardisation."""
        xb, lead = self._as_batch(x)
        y, ld = self._pass(FORWARD, 0, xb, xb.shape[0])
        return y.reshape(*lead, -1), ld.reshape(lead)

    def inverse(self, x, condition=None):
        """(x, log_det) of the inverse map (log_det = -sum s)."""
        xb, lead = self._as_batch(x)
        y, ld = self._pass(INVERSE, 0, xb, xb.shape[0])
        return y.reshape(*lead, -1), ld.reshape(lead)

    # ---------------------------------------------------------- training --
    def _workspace(self, B: int) -> _Workspace:
        ws = self._ws.get(B)
        if ws is None:
            ws = self._ws[B] = _Workspace(self, B)
        return ws

    def _backward(self, ws: _Workspace, B: int):
        s = self._struct()
        st = _native.stream_ptr()
        ws.grad.zero_()
        gy, gx = ws.gy, ws.gx
        for layer in reversed(range(self.n_layers)):
            _native.call("fmc_rnvp_backward_layer", st, C.byref(s), layer, ws.saved[layer].data_ptr(), B,
                         ws.g_logp.data_ptr(), gy.data_ptr(), gx.data_ptr(), ws.part.data_ptr(), ws.n_tiles)
            _native.call("fmc_rnvp_weight_grad", st, C.byref(s), layer, ws.part.data_ptr(), ws.n_tiles,
                         ws.part.shape[0], ws.grad.data_ptr())
            gy, gx = gx, gy
        return gy  # dL/d(input of layer 0)

    def grad_log_prob_sum(self, x, g_logp=None):
        """Hand-written adjoint: d/dtheta of sum_i g_logp[i] * log_prob(x_i) (flat,
        theta layout) and d/dx (standardised-space gradient divided by sigma)."""
        xb, _ = self._as_batch(x)
        B = xb.shape[0]
        ws = self._workspace(B)
        if g_logp is not None:
            ws.g_logp.copy_(torch.as_tensor(g_logp, dtype=torch.float32, device=self.device).reshape(B))
        else:
            ws.g_logp.fill_(1.0)
        self._pass(FORWARD, STANDARDIZE | BASE_LOGP, xb, B, saved=ws.saved, g_logp=ws.g_logp, gy_out=ws.gy,
                   want_x=False)
        gx = self._backward(ws, B)
        gx = gx / torch.sqrt(torch.diagonal(self.data_cov))
        out = ws.grad.clone(), gx.clone()
        ws.g_logp.fill_(-1.0 / B)
        return out

    def train_step(self, x, optim, state):
        """nf_model/base.py:102-125: loss, hand-written gradient, clip + AdamW (in place)."""
        xb, _ = self._as_batch(x)
        B = xb.shape[0]
        ws = self._workspace(B)
        self._ensure_packed()
        s = self._struct()
        _native.call("fmc_rnvp_pass", _native.stream_ptr(), C.byref(s), FORWARD, STANDARDIZE | BASE_LOGP,
                     xb.data_ptr(), B, 0, 0, 0, None, ws.logp.data_ptr(), ws.saved.data_ptr(), ws.g_logp.data_ptr(),
                     ws.gy.data_ptr())
        loss = -ws.logp.mean()
        self._backward(ws, B)
        state.count += 1
        mx = optim.max_norm if optim.max_norm is not None else 3.0e38
        _native.call("fmc_adamw_clip", _native.stream_ptr(), self.theta.data_ptr(), ws.grad.data_ptr(),
                     state.mu.data_ptr(), state.nu.data_ptr(), self.theta.numel(), float(optim.learning_rate),
                     float(optim.b1), float(optim.b2), float(optim.eps), float(optim.weight_decay), float(mx),
                     state.count, ws.norm_ws.data_ptr(), 1024, ws.gnorm.data_ptr())
        self._dirty = True
        return loss, self, state

    def save_resource(self, path):
        return self.save_model(path)

    def load_resource(self, path):
        return self.load_model(path)


class RealNVP(_RnvpBase):
    """RealNVP (realNVP.py:102-228): ``n_layers`` masked affine couplings, each
    with scale and shift MLPs [n_features, n_hidden, n_features] (ReLU).

    log_prob standardises with the data statistics and adds the log-density
    of N(0, I) — not of ``base_dist``, as in the reference (realNVP.py:211-213);
    ``sample`` draws from ``base_dist``.
    """

    def __repr__(self):
        return "RealNVP with n_features=" + str(self._n_features) + ", n_layers=" + str(self.n_layers)

    def __init__(self, n_features: int, n_layers: int, n_hidden: int, key, device=None, **kwargs):
        if getattr(kwargs.get("base_dist"), "learnable", False):
            # a trained covariance leaves the diagonal, which these kernels' base needs
            raise NotImplementedError("learnable base distributions are not supported by the RealNVP kernels "
                                      "(MaskedCouplingRQSpline trains them)")
        _native.require_gpu()
        self.device = torch.device("cuda") if device is None else torch.device(device)
        self._n_features = int(n_features)
        self.n_layers = int(n_layers)
        self.n_hidden = int(n_hidden)
        self.dt = 1.0
        self._masks = [realnvp_mask(self._n_features, i) for i in range(self.n_layers)]
        self._configure()
        theta = kwargs.get("_theta")
        if theta is None:
            theta = self._init_theta(key, 1e-4, kwargs)
        self.theta = theta.to(self.device, torch.float32).contiguous()
        if self.theta.numel() != self._theta_size:
            raise ValueError(f"theta has {self.theta.numel()} entries, the layout needs {self._theta_size}")
        self._dirty = True
        self._ws: dict = {}

    def log_prob(self, x):
        """realNVP.py:207-214. Differentiable in x: the backward pass runs the
        adjoint kernels (fmc_rnvp_backward_layer)."""
        if isinstance(x, torch.Tensor) and x.requires_grad and torch.is_grad_enabled():
            return _RnvpLogProb.apply(x, self)
        return self._log_prob_value(x)

    def _log_prob_value(self, x):
        xb, lead = self._as_batch(x)
        _, lp = self._pass(FORWARD, STANDARDIZE | BASE_LOGP, xb, xb.shape[0], want_x=False)
        return lp.reshape(lead)

    def sample(self, rng_key, n_samples: int):
        """realNVP.py:200-205, base noise from the device Philox stream of ``rng_key``."""
        k = frandom.key_words(rng_key)
        x, _ = self._pass(INVERSE, SAMPLE_BASE | STANDARDIZE, None, int(n_samples), key=k)
        return x

    def sample_from_noise(self, noise):
        """``sample`` with injected standard-normal base noise [n, d]."""
        xb, lead = self._as_batch(noise)
        x, _ = self._pass(INVERSE, SAMPLE_BASE | STANDARDIZE, xb, xb.shape[0])
        return x.reshape(*lead, -1)

    # --------------------------------------------------------------- I/O --
    def _config(self):
        return {"n_features": self._n_features, "n_layers": self.n_layers, "n_hidden": self.n_hidden}

    def save_model(self, path: str):
        np.savez(path + ".npz", theta=self.theta.detach().cpu().numpy(), config=json.dumps(self._config()))

    def load_model(self, path: str) -> "RealNVP":
        p = path if path.endswith(".npz") else path + ".npz"
        f = np.load(p)
        cfg = json.loads(str(f["config"]))
        if cfg != self._config():
            raise ValueError("saved flow has a different shape")
        return self.clone(torch.as_tensor(f["theta"], device=self.device))

    def print_parameters(self):
        print("RealNVP parameters:")
        print(f"Data mean: {self.data_mean}")
        print(f"Data covariance: {self.data_cov}")


class AffineCoupling(_RnvpBase):
    """AffineCoupling (realNVP.py:18-99): one coupling layer with scale and
    translate MLPs [n_features, n_hidden, n_features]; here ``mask`` = 1 marks
    the *transformed* dims (the MLPs read x * (1 - mask)), s = tanh(.) dt,
    y = (x + t) e^s, log_det = sum s. ``scale`` is the hidden-weight variance
    factor of the MLPs."""

    def __init__(self, n_features: int, n_hidden: int, mask, key, dt: float = 1, scale: float = 1e-4, device=None):
        _native.require_gpu()
        self.device = torch.device("cuda") if device is None else torch.device(device)
        self._n_features = int(n_features)
        self.n_layers = 1
        self.n_hidden = int(n_hidden)
        self.dt = float(dt)
        self._mask = torch.as_tensor(mask, dtype=torch.float32).reshape(-1).cpu()
        self._masks = [1 - self._mask]  # conditioning = 1 in the kernels' convention
        self._configure()
        self.theta = self._init_theta(key, scale, {}).to(self.device).contiguous()
        self._dirty = True
        self._ws: dict = {}

    @property
    def mask(self):
        return self._mask

    def __call__(self, x):
        return self.forward(x)

    def log_prob(self, x):
        raise NotImplementedError

    def sample(self, rng_key, n_samples: int):
        raise NotImplementedError

    def print_parameters(self):
        raise NotImplementedError

    def save_model(self, path: str):
        raise NotImplementedError

    def load_model(self, path: str):
        raise NotImplementedError


class _RnvpLogProb(torch.autograd.Function):
    """log_prob(x) with d/dx from the adjoint kernels."""

    @staticmethod
    def forward(ctx, x, model):
        ctx.model = model
        ctx.save_for_backward(x)
        return model._log_prob_value(x.detach()).to(x.dtype)

    @staticmethod
    def backward(ctx, g):
        (x,) = ctx.saved_tensors
        model = ctx.model
        _, gx = model.grad_log_prob_sum(x.detach().reshape(-1, model.n_features), g.reshape(-1).float())
        return gx.reshape(x.shape).to(g.dtype), None
