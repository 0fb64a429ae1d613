"""MaskedCouplingRQSpline — src/flowMC/resource/model/nf_model/rqSpline.py:364-507.

All arithmetic runs in the flow kernels of libflowmc_hip (fmc_flow_pass,
fmc_flow_backward_layer, fmc_flow_weight_grad, fmc_adamw_clip). Parameters
live in one flat FP32 master vector ``theta`` (layout documented in
flowmc_amd/csrc/fmc_flow_layout.h); the conditioner GEMMs read an MFMA
operand image packed from it in ``precision`` ("bfloat16" default, or
"float32") after every update.
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
from ..common import MLP, Gaussian, Linear, MaskedCouplingLayer, ScalarAffine
from .base import NFModel

FORWARD, INVERSE, INVERSE_EXACT = 0, 1, 2
STANDARDIZE, BASE_LOGP, SAMPLE_BASE = 1, 2, 4
_PREC = {"float32": 0, "fp32": 0, "bfloat16": 1, "bf16": 1}


class _Workspace:
    """Device scratch of one training batch size (allocated once, reused)."""

    def __init__(self, model: "MaskedCouplingRQSpline", B: int):
        dev = model.theta.device
        d, L = model.n_features, model.n_layers
        self.B = B
        self.ld = (B + 63) // 64 * 64
        tdt = torch.bfloat16 if model.precision == "bfloat16" else torch.float32
        self.saved = torch.empty((L, B, d), dtype=torch.float32, device=dev)
        self.logp = torch.empty(B, dtype=torch.float32, device=dev)
        self.g_logp = torch.full((B,), -1.0 / B, dtype=torch.float32, device=dev)
        self.gy = torch.empty((B, d), dtype=torch.float32, device=dev)
        self.gx = torch.empty((B, d), dtype=torch.float32, device=dev)
        # The weight gradient of layer l runs on a side stream beside the
        # adjoint of layer l-1: the split product's 48 KB workgroups fit on a
        # CU beside an adjoint workgroup (68 KB of LDS at d=64, [256, 256]).
        # Measured per training step at batch 10k: d=64, 10 layers, [256, 256]
        # 1.14 ms on one stream -> 1.02 ms overlapped; d=32, 8 layers,
        # [128, 128] 0.554 ms (per-tile kernel overlapped) -> 0.527 ms. Overlap
        # needs two stores and affine-partial buffers: the weight gradient
        # reads one while the next adjoint writes the other.
        self.overlap = True
        n_buf = 2 if self.overlap else 1
        self.store = [torch.empty((model._store_rows, self.ld), dtype=tdt, device=dev) for _ in range(n_buf)]
        self.n_part = self.ld // model._tile_bwd
        self.aff = [torch.empty((self.n_part, 2), dtype=torch.float32, device=dev) for _ in range(n_buf)]
        if self.overlap:
            self.side = torch.cuda.Stream(device=dev)  # default priority
        self.grad = torch.zeros(model.theta.numel(), dtype=torch.float32, device=dev)
        # partial blocks of the split weight-gradient product (fmc_flow_weight_grad_split)
        nbytes = C.c_int64(0)
        _native.call("fmc_flow_weight_grad_ws_size", C.byref(model._struct()), self.ld, C.byref(nbytes))
        self.wg_ws = torch.empty(max(1, nbytes.value // 4), dtype=torch.float32, device=dev)
        self.norm_ws = torch.empty(1024, dtype=torch.float64, device=dev)
        self.gnorm = torch.empty(1, dtype=torch.float32, device=dev)


class MaskedCouplingRQSpline(NFModel):
    # weight gradients by the split-K product over the batch (one workgroup
    # of 48 KB of LDS per block and batch range, fmc_flow_weight_grad_split)
    # or by one workgroup per 32 x 32 tile (fmc_flow_weight_grad); equal to
    # rounding. None: the split product (overlapped with the next adjoint,
    # batch 10k: d=64, [256, 256] 1.02 vs 1.25 ms per training step with the
    # per-tile kernel; d=32, [128, 128] 0.527 vs 0.557 ms)
    wgrad_split = None

    def __repr__(self):
        return "MaskedCouplingRQSpline with n_features=" + str(self._n_features) + ", n_layers=" + str(
            self.n_layers)

    def __init__(self, n_features: int, n_layers: int, hidden_size: list, num_bins: int, key,
                 spline_range=(-10.0, 10.0), precision: str = "bfloat16", device=None, **kwargs):
        if not 1 <= len(hidden_size) <= _native.FLOW_MAX_HIDDEN:
            raise NotImplementedError(f"the flow kernels implement conditioners with 1 to "
                                      f"{_native.FLOW_MAX_HIDDEN} hidden layers (got hidden_size={list(hidden_size)})")
        if precision not in _PREC:
            raise ValueError(f"precision must be one of {list(_PREC)}")
        _native.require_gpu()
        self.device = torch.device("cuda") if device is None else torch.device(device)
        self._n_features = int(n_features)
        self.n_layers = int(n_layers)
        self.hidden_size = [int(h) for h in hidden_size]
        self.num_bins = int(num_bins)
        self.spline_range = (float(spline_range[0]), float(spline_range[1]))
        self.min_bin_size = 1e-4
        self.min_knot_slope = 1e-4
        self.precision = "bfloat16" if _PREC[precision] == 1 else "float32"
        self.exact_inverse = bool(kwargs.get("exact_inverse", False))
        bd = kwargs.get("base_dist")
        # Gaussian(learnable=True) (common.py:243-273): the base mean and
        # covariance are trained with the flow
        self._base_learnable = bool(kwargs.get("_base_learnable", getattr(bd, "learnable", False)))
        theta = kwargs.get("_theta")
        if theta is None:
            theta = self._init_theta(key, kwargs)
        self.theta = theta.to(self.device, torch.float32).contiguous()
        # a base covariance with off-diagonal entries takes the full-covariance
        # base Gaussian (its Cholesky factor is packed with the biases)
        bc = self.theta[-self._n_features ** 2:].view(self._n_features, self._n_features)
        self._base_dense = bool(torch.count_nonzero(bc - torch.diag(torch.diagonal(bc))) > 0)
        self._base_dense = self._base_dense or self._base_learnable  # its covariance leaves the diagonal
        self._configure()
        self._dirty = True
        self._ws: dict[int, _Workspace] = {}
        self._slots: dict = {}  # train()'s working copies, shared with clones like _ws

    # ------------------------------------------------------------ layout --
    def _configure(self):
        probe = self._struct(theta_ptr=1)
        th, pk, pb, rows = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        tf, tb = C.c_int32(), C.c_int32()
        _native.call("fmc_flow_sizes", C.byref(probe), C.byref(th), C.byref(pk), C.byref(pb), C.byref(rows),
                     C.byref(tf), C.byref(tb))
        self._theta_size, self._packed_size, self._pbias_size = th.value, pk.value, pb.value
        self._store_rows, self._tile_fwd, self._tile_bwd = rows.value, tf.value, tb.value
        if self._tile_fwd == 0 or self._tile_bwd == 0:
            raise NotImplementedError("flow dimensions exceed the kernels' LDS budget")
        tdt = torch.bfloat16 if self.precision == "bfloat16" else torch.float32
        self.packed = torch.empty(self._packed_size, dtype=tdt, device=self.device)
        self.packed_bias = torch.empty(self._pbias_size, dtype=torch.float32, device=self.device)
        self._layer_size = 2 + sum(int(np.prod(shp)) for _, shp in self._leaf_shapes()[2:])
        d = self._n_features
        off = self.n_layers * self._layer_size
        self._off = {"data_mean": off, "data_cov": off + d, "base_mean": off + d + d * d,
                     "base_cov": off + 2 * d + d * d}
        assert off + 2 * d + 2 * d * d == self._theta_size

    def _struct(self, theta_ptr=None) -> _native.fmc_flow:
        tp = theta_ptr if theta_ptr is not None else _native.dptr(self.theta, name="theta")
        pk = getattr(self, "packed", None)
        pb = getattr(self, "packed_bias", None)
        return _native.make_flow(self._n_features, self.n_layers, self.num_bins, self.hidden_size,
                                 _PREC[self.precision], self.spline_range[0], self.spline_range[1], self.min_bin_size,
                                 self.min_knot_slope, tp,
                                 pk.data_ptr() if (pk is not None and theta_ptr is None) else None,
                                 pb.data_ptr() if (pb is not None and theta_ptr is None) else None,
                                 flags=_native.FLOW_DENSE_BASE if getattr(self, "_base_dense", False) else 0)

    def _leaf_shapes(self):
        """(name, shape) of one layer's leaves in theta order: scale, shift,
        then W{i}, b{i} of the conditioner's linear layers (W1 .. Wn hidden,
        W{n+1} the output layer; eqx.nn.Linear's [out, in] weights)."""
        d, P = self._n_features, 3 * self.num_bins + 1
        ins = [d] + self.hidden_size
        outs = self.hidden_size + [d * P]
        shapes = [("scale", ()), ("shift", ())]
        for i, (a, b) in enumerate(zip(ins, outs)):
            shapes += [(f"W{i + 1}", (b, a)), (f"b{i + 1}", (b,))]
        return shapes

    def view(self, layer: int, name: str) -> torch.Tensor:
        """View of one leaf of layer ``layer`` inside theta."""
        o = layer * self._layer_size
        for k, shp in self._leaf_shapes():
            n = int(np.prod(shp)) if shp else 1
            if k == name:
                return self.theta[o:o + n].view(shp)
            o += n
        raise KeyError(name)

    def _stat(self, name):
        d = self._n_features
        o = self._off[name]
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
        return Gaussian(self._stat("base_mean"), self._stat("base_cov"), learnable=self._base_learnable)

    @property
    def layers(self):
        """Reference-shaped views: [MaskedCouplingLayer(ScalarAffine), MaskedCouplingLayer(RQSpline MLP)] per layer."""
        out = []
        d = self._n_features
        for i in range(self.n_layers):
            aff = MaskedCouplingLayer(ScalarAffine(self.view(i, "scale"), self.view(i, "shift")),
                                      torch.zeros(d, dtype=torch.bool))
            mlp = MLP([Linear(self.view(i, f"W{k}"), self.view(i, f"b{k}"))
                       for k in range(1, len(self.hidden_size) + 2)])
            mask = ((torch.arange(d) + i) % 2).to(torch.bool)
            out.append([aff, MaskedCouplingLayer(mlp, mask)])
        return out

    # -------------------------------------------------------------- init --
    def _init_theta(self, key, kwargs) -> torch.Tensor:
        """Initialisation with the distributions of rqSpline.py:427-443 / common.py:84-107:
        hidden weights N(0, 1e-2 / fan_in), hidden biases and the last layer
        U(-1/sqrt(fan_in), 1/sqrt(fan_in)), ScalarAffine (0, 0). The draws come
        from a torch generator seeded from ``key`` (not JAX's streams)."""
        d = self._n_features
        k0, k1 = frandom.key_words(key)
        g = torch.Generator().manual_seed((k0 << 32) | k1)
        n_lin = len(self.hidden_size) + 1
        parts = []
        for _ in range(self.n_layers):
            parts += [torch.zeros(2)]
            fan = d
            for k, shp in self._leaf_shapes()[2:]:
                if k[0] == "W":  # weights come before their biases: fan is W's fan-in for both
                    fan = shp[1]
                    if int(k[1:]) == n_lin:
                        parts.append((torch.rand(*shp, generator=g) * 2 - 1) / math.sqrt(fan))
                    else:
                        parts.append(torch.randn(*shp, generator=g) * math.sqrt(1e-2 / fan))
                else:
                    parts.append((torch.rand(*shp, generator=g) * 2 - 1) / math.sqrt(fan))
        dm = kwargs.get("data_mean")
        dc = kwargs.get("data_cov")
        bd = kwargs.get("base_dist")
        dm = torch.zeros(d) if dm is None else torch.as_tensor(dm, dtype=torch.float32).reshape(d)
        dc = torch.eye(d) if dc is None else torch.as_tensor(dc, dtype=torch.float32).reshape(d, d)
        if bd is None:
            bm, bc = torch.zeros(d), torch.eye(d)
        else:
            bm = torch.as_tensor(bd.mean, dtype=torch.float32).reshape(d).cpu()
            bc = torch.as_tensor(bd.cov, dtype=torch.float32).reshape(d, d).cpu()
        parts += [dm.reshape(-1).cpu(), dc.reshape(-1).cpu(), bm.reshape(-1), bc.reshape(-1)]
        return torch.cat([p.reshape(-1).to(torch.float32) for p in parts])

    def clone(self, theta: Optional[torch.Tensor] = None) -> "MaskedCouplingRQSpline":
        m = MaskedCouplingRQSpline.__new__(MaskedCouplingRQSpline)
        m.__dict__.update({k: v for k, v in self.__dict__.items() if k not in ("theta", "packed", "packed_bias", "_ws")})
        m._configure()
        m.theta = (self.theta if theta is None else theta).clone()
        m._dirty = True
        # the per-batch-size workspaces (scratch of a training step or
        # gradient call, sized by the flow's shape and precision alone) are
        # shared with the clone: training clones the model every call
        # (nf_model/base.py:153-210) and would otherwise allocate ~100 MB of
        # scratch each time. Calls on the two models run in stream order.
        m._ws = self._ws
        return m

    def _training_copy(self, state):
        """The (model, optimiser state) pair that train() updates: the
        clone and state copy of nf_model/base.py:153-210, kept with the shared
        workspaces (one pair per precision and base kind) and refilled from
        self and state on every call. Every train call of this flow and its
        clones then updates the same parameter and moment buffers, so a
        recorded training step (graph_train) stays valid from one call to
        the next. The pair never leaves train(): it returns a fresh clone of
        the best epoch."""
        key = (self.precision, self._base_learnable, self._base_dense)
        slot = self._slots.get(key)
        if slot is None or slot[1].mu.shape != state.mu.shape:
            slot = self._slots[key] = (self.clone(), state.clone())
            return slot
        m, st = slot
        own = {k: m.__dict__[k] for k in ("theta", "packed", "packed_bias")}
        m.__dict__.update({k: v for k, v in self.__dict__.items() if k not in own})
        m.__dict__.update(own)
        m.theta.copy_(self.theta)
        m._dirty = True
        st.mu.copy_(state.mu)
        st.nu.copy_(state.nu)
        st.count = state.count
        return slot

    def set_data_stats(self, mean, cov):
        self.data_mean.copy_(mean)
        self.data_cov.copy_(cov)
        self._dirty = True

    def to_precision(self, precision: str = "float32") -> "MaskedCouplingRQSpline":
        """nf_model/base.py:212-242: precision of the conditioner GEMMs (theta stays FP32)."""
        p = precision.lower()
        if p not in _PREC:
            raise ValueError(f"Precision {precision} not supported. Choose from {list(_PREC)}")
        m = self.clone()
        m.precision = "bfloat16" if _PREC[p] == 1 else "float32"
        m._configure()
        m._ws = {}  # the stores hold the conditioner's precision
        m._slots = {}
        return m

    # -------------------------------------------------------------- kernels
    def _ensure_packed(self):
        if self._dirty:
            s = self._struct()
            _native.call("fmc_flow_pack", _native.stream_ptr(), C.byref(s))
            self._dirty = False

    def mark_updated(self):
        """Call after editing theta (or a view of it) in place."""
        self._dirty = True

    def _pass(self, direction, flags, x, n, key=(0, 0), saved=None, g_logp=None, gy_out=None, want_x=True):
        self._ensure_packed()
        d = self._n_features
        out = torch.empty((n, d), dtype=torch.float32, device=self.device) if want_x else None
        ld = torch.empty(n, dtype=torch.float32, device=self.device)
        s = self._struct()
        _native.call("fmc_flow_pass", _native.stream_ptr(), C.byref(s), direction, flags,
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
        """rqSpline.py:450-468: (y, log_det) with no standardisation."""
        xb, lead = self._as_batch(x)
        y, ld = self._pass(FORWARD, 0, xb, xb.shape[0])
        return y.reshape(*lead, -1), ld.reshape(lead)

    def inverse(self, x, condition=None):
        """rqSpline.py:470-488 (reference layer order unless exact_inverse)."""
        xb, lead = self._as_batch(x)
        y, ld = self._pass(INVERSE_EXACT if self.exact_inverse else INVERSE, 0, xb, xb.shape[0])
        return y.reshape(*lead, -1), ld.reshape(lead)

    def log_prob(self, x):
        """rqSpline.py:498-504. Differentiable in x: the backward pass runs the
        hand-written adjoint kernels (fmc_flow_backward_layer)."""
        if isinstance(x, torch.Tensor) and x.requires_grad and torch.is_grad_enabled():
            return _FlowLogProb.apply(x, self)
        return self._log_prob_value(x)

    def _log_prob_value(self, x):
        xb, lead = self._as_batch(x)
        _, lp = self._pass(FORWARD, STANDARDIZE | BASE_LOGP, xb, xb.shape[0], want_x=False)
        return lp.reshape(lead)

    def sample(self, rng_key, n_samples: int):
        """rqSpline.py:490-496, base noise from the device Philox stream of ``rng_key``."""
        k = frandom.key_words(rng_key)
        x, _ = self._pass(INVERSE_EXACT if self.exact_inverse else INVERSE, SAMPLE_BASE | STANDARDIZE, None,
                          int(n_samples), key=k)
        return x

    def sample_from_noise(self, noise):
        """``sample`` with injected standard-normal base noise [n, d]."""
        xb, lead = self._as_batch(noise)
        x, _ = self._pass(INVERSE_EXACT if self.exact_inverse else INVERSE, SAMPLE_BASE | STANDARDIZE, xb,
                          xb.shape[0])
        return x.reshape(*lead, -1)

    def sample_and_log_prob(self, rng_key=None, n_samples: int = 0, noise=None):
        """Samples and their flow log-density (``log_prob``'s value) from one
        exact-inverse pass: log N(z) minus the inverse log-determinant. Only
        with ``exact_inverse`` (the reference's inverse order is not the
        inverse of the forward map, so its samples do not have that density)."""
        if not self.exact_inverse:
            raise ValueError("sample_and_log_prob needs exact_inverse=True")
        if noise is not None:
            xb, lead = self._as_batch(noise)
            x, lq = self._pass(INVERSE_EXACT, SAMPLE_BASE | STANDARDIZE | BASE_LOGP, xb, xb.shape[0])
            return x.reshape(*lead, -1), lq.reshape(lead)
        k = frandom.key_words(rng_key)
        return self._pass(INVERSE_EXACT, SAMPLE_BASE | STANDARDIZE | BASE_LOGP, None, int(n_samples), key=k)

    # ------------------------------------------------------------ training
    def _workspace(self, B: int) -> _Workspace:
        ws = self._ws.get(B)
        if ws is None:
            ws = self._ws[B] = _Workspace(self, B)
        return ws

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
        self._ensure_packed()
        self._pass(FORWARD, STANDARDIZE | BASE_LOGP, xb, B, saved=ws.saved, g_logp=ws.g_logp, gy_out=ws.gy,
                   want_x=False)
        base_g = self._base_grad(ws) if self._base_learnable else None
        self._backward(ws, B)
        self._add_base_grad(ws, base_g)
        gx = ws.gy if self.n_layers % 2 == 0 else ws.gx
        gx = gx / torch.sqrt(torch.diagonal(self.data_cov))
        out = ws.grad.clone(), gx.clone()
        ws.g_logp.fill_(-1.0 / B)
        return out

    def _base_grad(self, ws: _Workspace):
        """Gradient of sum_n g_n log N(y_n; mean, cov) with respect to a
        learnable base mean and covariance (Gaussian.log_prob, common.py:285-286:
        jax.scipy.stats.multivariate_normal.logpdf, whose Cholesky factorisation
        symmetrises its input). The forward pass leaves gy_n = -g_n w_n with
        w_n = cov^-1 (y_n - mean) in ws.gy, so
            d/dmean = sum_n g_n w_n = -sum_n gy_n,
            d/dcov  = 1/2 sum_n g_n w_n w_n^T - 1/2 (sum_n g_n) cov^-1,
        with g_n w_n w_n^T = gy_n gy_n^T / g_n (0 where g_n = 0). Read before
        the adjoint reuses ws.gy; small device GEMMs."""
        g = ws.g_logp
        gy = ws.gy
        gm = -gy.sum(0)
        ig = torch.where(g != 0, 1.0 / torch.where(g != 0, g, torch.ones_like(g)), torch.zeros_like(g))
        S = gy.t().double() @ (gy.double() * ig.double()[:, None])
        cov = self._stat("base_cov").double()
        cinv = torch.cholesky_inverse(torch.linalg.cholesky(0.5 * (cov + cov.t())))
        gc = 0.5 * S - 0.5 * g.double().sum() * cinv
        return gm, gc.float()

    def _add_base_grad(self, ws: _Workspace, base_g):
        if base_g is None:
            return
        d = self._n_features
        om, oc = self._off["base_mean"], self._off["base_cov"]
        ws.grad[om:om + d] += base_g[0]
        ws.grad[oc:oc + d * d] += base_g[1].reshape(-1)

    def _backward(self, ws: _Workspace, B: int):
        """Per layer, last to first: the adjoint, then the weight gradient of
        that layer, in one library call (fmc_flow_backward). When the
        workspace overlaps them, the weight gradient runs on the side stream
        (it only reads its layer's store and affine partials) beside the next
        adjoint on the current stream; the stores alternate, an adjoint waits
        until the weight gradient that last read its store has finished, and
        the current stream joins the side stream before returning. The call
        clears ws.grad itself (beside the first adjoint when overlapping).
        Results do not depend on the schedule."""
        s = self._struct()
        st = torch.cuda.current_stream(self.device).cuda_stream
        split = self.wgrad_split if self.wgrad_split is not None else True
        two = ws.overlap and len(ws.store) > 1
        _native.call("fmc_flow_backward", st, ws.side.cuda_stream if two else None, C.byref(s), ws.saved.data_ptr(),
                     B, ws.g_logp.data_ptr(), ws.gy.data_ptr(), ws.gx.data_ptr(), ws.store[0].data_ptr(),
                     ws.store[1].data_ptr() if two else None, ws.ld, ws.aff[0].data_ptr(),
                     ws.aff[1].data_ptr() if two else None, ws.n_part, ws.grad.data_ptr(), int(split),
                     ws.wg_ws.data_ptr(), ws.wg_ws.numel() * 4)

    # True: training steps of one batch size, parameter buffers, optimiser
    # state and hyperparameters are recorded once as a HIP graph (pack,
    # forward, the two-stream adjoint / weight-gradient schedule, clip +
    # AdamW with the step count in device memory) and replayed, bit for bit
    # the eager steps. A recording costs ~2 ms and saves ~16-30 us a step
    # (batch 10k, d=32: 0.510 -> 0.492 ms per step). train() updates the
    # same working buffers on every call (_training_copy), so one recording
    # serves every training round of a sampling run.
    graph_train = True

    def train_step(self, x, optim, state):
        """nf_model/base.py:102-125: loss, hand-written gradient, clip + AdamW (in place)."""
        xb, _ = self._as_batch(x)
        B = xb.shape[0]
        ws = self._workspace(B)
        if self.graph_train and not self._base_learnable:
            return self._train_step_graph(xb, ws, optim, state)
        self._ensure_packed()
        loss = self._train_body(xb, ws, optim, state, None)
        self._dirty = True
        return loss, self, state

    def _train_step_no_loss(self, x, optim, state):
        """train_step without the loss reduction: the same parameter update,
        for the steps of an epoch whose loss nobody reads."""
        xb, _ = self._as_batch(x)
        ws = self._workspace(xb.shape[0])
        if self.graph_train and not self._base_learnable:
            _, model, state = self._train_step_graph(xb, ws, optim, state, want_loss=False)
            return model, state
        self._ensure_packed()
        self._train_body(xb, ws, optim, state, None, want_loss=False)
        self._dirty = True
        return self, state

    def _train_body(self, xb, ws, optim, state, step_dev, want_loss=True):
        """Forward, loss, adjoint, clip + AdamW of one batch; with step_dev the
        AdamW step count is read from device memory (the graph body, which
        leaves the loss to the caller)."""
        B = xb.shape[0]
        s = self._struct()
        _native.call("fmc_flow_pass", _native.stream_ptr(), C.byref(s), FORWARD, STANDARDIZE | BASE_LOGP,
                     xb.data_ptr(), B, 0, 0, 0, None, ws.logp.data_ptr(), ws.saved.data_ptr(), ws.g_logp.data_ptr(),
                     ws.gy.data_ptr())
        loss = -ws.logp.mean() if want_loss else None
        base_g = self._base_grad(ws) if self._base_learnable else None
        self._backward(ws, B)
        self._add_base_grad(ws, base_g)
        mx = optim.max_norm if optim.max_norm is not None else 3.0e38
        hyper = (float(optim.learning_rate), float(optim.b1), float(optim.b2), float(optim.eps),
                 float(optim.weight_decay), float(mx))
        if step_dev is None:
            state.count += 1
            _native.call("fmc_adamw_clip", _native.stream_ptr(), self.theta.data_ptr(), ws.grad.data_ptr(),
                         state.mu.data_ptr(), state.nu.data_ptr(), self.theta.numel(), *hyper, state.count,
                         ws.norm_ws.data_ptr(), 1024, ws.gnorm.data_ptr())
        else:
            _native.call("fmc_adamw_clip_dev", _native.stream_ptr(), self.theta.data_ptr(), ws.grad.data_ptr(),
                         state.mu.data_ptr(), state.nu.data_ptr(), self.theta.numel(), *hyper, step_dev.data_ptr(),
                         ws.norm_ws.data_ptr(), 1024, ws.gnorm.data_ptr())
        return loss

    def _train_step_graph(self, xb, ws, optim, state, want_loss=True):
        """One replay of the recorded step; the loss (read from the forward's
        log-densities, which the rest of the step leaves alone) only when
        asked for."""
        B = xb.shape[0]
        # every address and value the recorded launches carry: a replay is
        # only valid for the same buffers (the workspace is shared by clones)
        key = (B, self.theta.data_ptr(), self.packed.data_ptr(), self.packed_bias.data_ptr(), state.mu.data_ptr(),
               state.nu.data_ptr(), float(optim.learning_rate), float(optim.b1), float(optim.b2), float(optim.eps),
               float(optim.weight_decay), optim.max_norm, self._base_dense)
        if getattr(ws, "graph_key", None) != key:
            ws.graph = None
            ws.x_static = torch.empty((B, self._n_features), dtype=torch.float32, device=self.device)
            ws.step_dev = torch.zeros(1, dtype=torch.int32, device=self.device)
            ws.step_next = None  # the count step_dev holds for the next replay
            self._ensure_packed()  # (recording runs nothing: theta stays as it is)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                s = self._struct()
                _native.call("fmc_flow_pack", _native.stream_ptr(), C.byref(s))
                self._train_body(ws.x_static, ws, optim, state, ws.step_dev, want_loss=False)
                ws.step_dev.add_(1)  # the next replay's count, without a launch outside the graph
            ws.graph, ws.graph_key = g, key
        ws.x_static.copy_(xb)
        state.count += 1
        if ws.step_next != state.count:  # first replay, or a state whose count moved elsewhere
            ws.step_dev.fill_(state.count)
        ws.graph.replay()
        ws.step_next = state.count + 1
        self._dirty = True
        return (-ws.logp.mean() if want_loss else None), self, state

    # --------------------------------------------------------------- I/O --
    def _config(self):
        return {"n_features": self._n_features, "n_layers": self.n_layers, "hidden_size": self.hidden_size,
                "num_bins": self.num_bins, "spline_range": list(self.spline_range), "precision": self.precision,
                "learnable_base": self._base_learnable, "dense_base": self._base_dense}

    def save_model(self, path: str):
        np.savez(path + ".npz", theta=self.theta.detach().cpu().numpy(), config=json.dumps(self._config()))

    def load_model(self, path: str) -> "MaskedCouplingRQSpline":
        p = path if path.endswith(".npz") else path + ".npz"
        f = np.load(p)
        cfg = json.loads(str(f["config"]))
        if cfg["n_features"] != self._n_features or cfg["n_layers"] != self.n_layers:
            raise ValueError("saved flow has a different shape")
        m = self.clone(torch.as_tensor(f["theta"], device=self.device))
        # the saved flow's base path (full covariance or diagonal, learnable
        # or not); a file without it takes the path its covariance implies
        m._base_learnable = bool(cfg.get("learnable_base", False))
        if "dense_base" in cfg:
            m._base_dense = bool(cfg["dense_base"])
        else:
            bc = m._stat("base_cov")
            m._base_dense = bool(torch.count_nonzero(bc - torch.diag(torch.diagonal(bc))) > 0) or m._base_learnable
        m._configure()  # the packed-bias image holds the Cholesky factor only on the full-covariance path
        m._dirty = True
        m._ws = {}
        m._slots = {}
        return m

    def print_parameters(self):
        print("RQSpline parameters:")
        print(self._config())


class _FlowLogProb(torch.autograd.Function):
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
