"""Many independent HMC problems of uneven dimension advanced in ONE launch.

Each problem ("slot") is a built-in target with its own dimension, chains,
position / log-prob / acceptance buffers and lanes-per-chain; the work
items (slot, first chain, n chains) map one workgroup each. Per chain the
step is HMC.kernel (src/flowMC/resource/kernel/HMC.py:98-151) exactly as
in the single-target fused kernel (fmc_hmc_run); this is the batched
form of TakeSerialSteps(HMC) over several LogPDFs.
"""

from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import torch

from ... import _native
from ... import autotune
from ... import random as frandom
from ...targets import Target
from .HMC import condition_spec


def _pow2ceil(v: int) -> int:
    p = 1
    while p < v:
        p *= 2
    return p


class BatchedHMC:
    def __init__(self, targets: Sequence[Target], n_chains, step_size: float = 0.05, n_leapfrog: int = 10,
                 metrics: Optional[Sequence[Optional[torch.Tensor]]] = None, device="cuda"):
        _native.require_gpu()
        self.targets = list(targets)
        T = len(self.targets)
        self.n_chains = [int(n_chains)] * T if isinstance(n_chains, int) else [int(c) for c in n_chains]
        self.dims = [t.n_dim for t in self.targets]
        self.step_size = float(step_size)
        self.n_leapfrog = int(n_leapfrog)
        self.device = torch.device(device)
        # per slot: None (identity), a diagonal [d] or a condition matrix [d, d]
        self.metrics = list(metrics) if metrics is not None else [None] * T
        self.x = [torch.zeros((c, d), dtype=torch.float32, device=self.device) for c, d in zip(self.n_chains, self.dims)]
        self.logp = [torch.zeros(c, dtype=torch.float32, device=self.device) for c in self.n_chains]
        self.layout = None  # (vec_per_lane, block_size, lanes_per_chain per slot)
        self.specialise = True  # False: always the run-time dispatch kernel

    def init_positions(self, key):
        for i, x in enumerate(self.x):
            x.copy_(frandom.normal(frandom.fold_in(key, i), tuple(x.shape), device=self.device))

    # ------------------------------------------------------------- tables --
    def default_layout(self, nv: int = 1, block: int = 256, spread: int = 1):
        lpcs = [min(64, spread * _pow2ceil(max(1, -(-d // (4 * nv))))) for d in self.dims]
        for d, l in zip(self.dims, lpcs):
            if 4 * nv * l < d:
                raise ValueError(f"vec_per_lane={nv} cannot hold d={d} in 64 lanes")
        return nv, block, lpcs

    def uniform_kind(self) -> int:
        """The target kind every slot shares when no slot has a condition
        matrix (the launch then runs the kernel specialised on that kind and on
        each slot's lanes per chain), else -1 (the run-time dispatch kernel)."""
        if self.specialise and all(m is None for m in self.metrics):
            kinds = {t.kind for t in self.targets}
            if len(kinds) == 1:
                return int(kinds.pop())
        return -1

    def _metric(self, s):
        """Slot s's (kinetic weights, momentum factor): a [d] vector is taken
        as the diagonal as given (already on the device); a [d, d] matrix goes
        through HMC's condition_spec (dense: colsum(M), chol(inv(M)))."""
        m = self.metrics[s]
        if m is None or m.ndim == 1:
            return m, None
        cache = self.__dict__.setdefault("_specs", {})
        key = (s, id(m))
        if key not in cache:
            cache[key] = condition_spec(m, self.dims[s], self.device)
        return cache[key]

    def _check_slot(self, s, out, noise, log_u, n_steps, step0, thinning, n_run):
        """Host-side checks of slot ``s`` (the kernel reads the slot table from
        device memory, so nothing on the device side can validate it): every
        tensor the kernel will touch must have the shape the slot implies and
        every buffer row it writes must exist. Shapes are checked first, then
        dtype / device / contiguity (by the caller, once every slot's shapes
        have passed). Returns (buf_len, cursor, [(name, tensor)])."""
        t, C_, d = self.targets[s], self.n_chains[s], self.dims[s]
        # the descriptor's kind / dims / parameter count (host-side check only)
        desc = _native.fmc_target(t.kind, t.n_dim, t.n_modes, 0, t.params.data_ptr(), t._n_params)
        _native.call("fmc_target_check", C.byref(desc))
        if t.n_dim != d:
            raise ValueError(f"slot {s}: target n_dim {t.n_dim} != {d}")

        def shape_is(name, ten, want):
            got = tuple(ten.shape)
            if len(got) != len(want) or any(w is not None and g != w for g, w in zip(got, want)):
                txt = "[" + ", ".join("T" if w is None else str(w) for w in want) + "]"
                raise ValueError(f"slot {s}: {name} must be {txt}, got {list(got)}")

        shape_is("state", self.x[s], (C_, d))
        shape_is("log_prob state", self.logp[s], (C_,))
        m, chol = self._metric(s)
        if m is not None:
            shape_is("metric", m, (d,))
        if (noise is None) != (log_u is None):
            raise ValueError(f"slot {s}: noise and log_u are injected together")
        if noise is not None:
            shape_is("noise", noise, (n_steps, C_, d))
            shape_is("log_u", log_u, (n_steps, C_))
        pos, lps, accs, cursor = out if out is not None else (None, None, None, 0)
        bufs = [(n, b) for n, b in (("position buffer", pos), ("log_prob buffer", lps), ("acceptance buffer", accs))
                if b is not None]
        lens = set()
        for name, buf in bufs:
            shape_is(name, buf, (C_, None, d) if name.startswith("position") else (C_, None))
            lens.add(int(buf.shape[1]))
        if len(lens) > 1:
            raise ValueError(f"slot {s}: output buffers disagree on their length {sorted(lens)}")
        buf_len, cursor = (lens.pop(), int(cursor)) if lens else (1, 0)
        if bufs and n_run > 0:
            last_row = cursor + (step0 + n_run - 1) // thinning
            if cursor < 0 or last_row >= buf_len:
                raise ValueError(f"slot {s}: rows {cursor}..{last_row} do not fit a buffer of length {buf_len}")
        tensors = [(f"slot {s} {name}", ten) for name, ten in
                   [("state", self.x[s]), ("log_prob state", self.logp[s]), ("metric", m), ("momentum factor", chol),
                    ("noise", noise),
                    ("log_u", log_u)] + bufs]
        return buf_len, cursor, tensors

    def _tables(self, layout, key, n_steps_noise, outputs, noise, log_u, thinning=1, step0=0, n_run=None):
        nv, block, lpcs = layout
        if thinning < 1:
            raise ValueError("thinning must be >= 1")
        if outputs is not None and len(outputs) != len(self.targets):
            raise ValueError(f"outputs must have one entry per target ({len(self.targets)})")
        for d, l in zip(self.dims, lpcs):
            # a chain's lanes must sit in one wavefront: power of two <= 64
            if l < 1 or l > 64 or l & (l - 1) or block % l or 4 * nv * l < d:
                raise ValueError(f"lanes_per_chain={l} (vec_per_lane={nv}, block={block}) cannot hold d={d}")
        n_run = n_steps_noise if n_run is None else n_run
        checked = []
        for s in range(len(self.targets)):
            checked.append(self._check_slot(s, outputs[s] if outputs is not None else None,
                                            noise[s] if noise is not None else None,
                                            log_u[s] if log_u is not None else None,
                                            n_steps_noise, step0, thinning, n_run))
        for _, _, tensors in checked:
            for name, ten in tensors:
                _native.dptr(ten, name=name)
        slots = (_native.fmc_hmc_slot * len(self.targets))()
        items = []
        self._keep = []
        for s, t in enumerate(self.targets):
            k0, k1 = frandom.key_words(frandom.fold_in(key, s))
            out = outputs[s] if outputs is not None else None
            m, chol = self._metric(s)
            nz = noise[s] if noise is not None else None
            lu = log_u[s] if log_u is not None else None
            buf_len, cursor, _ = checked[s]
            out = out if out is not None else (None, None, None, 0)
            self._keep += [m, chol, nz, lu]
            slots[s].target = t.descriptor()
            slots[s].x = self.x[s].data_ptr()
            slots[s].logp = self.logp[s].data_ptr()
            slots[s].metric = _native.dptr(m, name="metric")
            slots[s].momentum_chol = _native.dptr(chol, name="momentum factor")
            slots[s].noise = _native.dptr(nz, name="noise")
            slots[s].log_u = _native.dptr(lu, name="log_u")
            slots[s].out_pos = _native.dptr(out[0], name="position buffer")
            slots[s].out_logp = _native.dptr(out[1], name="log_prob buffer")
            slots[s].out_acc = _native.dptr(out[2], name="acceptance buffer")
            slots[s].n_chains = self.n_chains[s]
            slots[s].buf_len = buf_len
            slots[s].cursor = cursor
            slots[s].key0, slots[s].key1 = k0, k1
            slots[s].lanes_per_chain = lpcs[s]
            cpb = block // lpcs[s]
            for c0 in range(0, self.n_chains[s], cpb):
                items.append((s, c0, min(cpb, self.n_chains[s] - c0), 0))
        # order work items by cost (dim) so long rows start first
        items.sort(key=lambda it: -self.dims[it[0]])
        sb = bytes(slots)
        slot_dev = torch.frombuffer(bytearray(sb), dtype=torch.uint8).to(self.device)
        it_dev = torch.tensor(items, dtype=torch.int32).reshape(-1).to(self.device)
        return slot_dev, it_dev, len(items)

    # --------------------------------------------------------------- run ---
    def run(self, key, n_steps: int, outputs=None, thinning: int = 1, noise=None, log_u=None, fused_steps: int = 16,
            layout=None):
        """Advance every problem by ``n_steps`` HMC steps. ``outputs[s]`` is
        (positions [C,T,d], log_probs [C,T], accepts [C,T], cursor) or None;
        ``noise[s]`` [n_steps, C, d] / ``log_u[s]`` [n_steps, C] inject randomness."""
        if noise is None and log_u is None:
            # the slot / work-item tables are the same for every launch of this
            # call (the step offset is a kernel argument): upload them once
            launch = self.prepare(key, n_steps, outputs, thinning, fused_steps, layout)
            launch()
            self._last_tables = launch.tables  # keep alive until the stream passes them
            return
        layout = layout or self.layout or self.default_layout()
        nv, block, lpcs = layout
        st = _native.stream_ptr()
        maxd = max(self.dims)
        for s0 in range(0, n_steps, fused_steps):
            # injected noise: each launch's slots point at its own noise window
            ns = min(fused_steps, n_steps - s0)
            nz = [n[s0:s0 + ns] for n in noise] if noise is not None else None
            lu = [u[s0:s0 + ns] for u in log_u] if log_u is not None else None
            slot_dev, it_dev, n_items = self._tables(layout, key, ns, outputs, nz, lu, thinning, s0, ns)
            _native.call("fmc_hmc_run_batched", st, slot_dev.data_ptr(), len(self.targets), it_dev.data_ptr(), n_items,
                         maxd, ns, s0, self.step_size, self.n_leapfrog, thinning, nv, block, self.uniform_kind())
            self._last_tables = (slot_dev, it_dev)  # keep alive until the stream passes them

    def prepare(self, key, n_steps: int, outputs=None, thinning: int = 1, fused_steps: int = 16, layout=None):
        """Upload the slot / work-item tables of ``run(key, n_steps, outputs,
        thinning, fused_steps=fused_steps)`` (no injected noise) and return a
        callable that issues its launches on the current stream. The callable
        allocates and copies nothing, so it can be recorded in a HIP graph
        (``torch.cuda.graph``) and replayed; each call or replay advances the
        chains from their current state with the same keys and step offsets."""
        layout = layout or self.layout or self.default_layout()
        nv, block, lpcs = layout
        maxd = max(self.dims)
        slot_dev, it_dev, n_items = self._tables(layout, key, fused_steps, outputs, None, None, thinning, 0, n_steps)
        n_slots, eps, n_lf, uk = len(self.targets), self.step_size, self.n_leapfrog, self.uniform_kind()

        def launch():
            st = _native.stream_ptr()
            for s0 in range(0, n_steps, fused_steps):
                ns = min(fused_steps, n_steps - s0)
                _native.call("fmc_hmc_run_batched", st, slot_dev.data_ptr(), n_slots, it_dev.data_ptr(), n_items,
                             maxd, ns, s0, eps, n_lf, thinning, nv, block, uk)

        launch.tables = (slot_dev, it_dev)
        return launch

    def tune(self, key, n_steps: int = 16):
        """Pick (vec_per_lane, block size, lane spread) by timing short runs
        (one launch of ``n_steps`` fused steps per candidate, averaged over
        the autotuner's repetitions)."""
        sig = ("hmc_batched", tuple(self.dims), tuple(self.n_chains), self.n_leapfrog, self.uniform_kind())
        saved = [x.clone() for x in self.x]
        cands = []
        for nv in (1, 2):
            for block in (128, 256, 512):
                for spread in (1, 2):
                    try:
                        lay = self.default_layout(nv, block, spread)
                    except ValueError:
                        continue
                    if lay not in cands:
                        cands.append(lay)

        launches = {}

        def go(layout):
            # tables built and uploaded once per candidate: only the launch is timed
            lk = (layout[0], layout[1], tuple(layout[2]))
            if lk not in launches:
                launches[lk] = self.prepare(key, n_steps, layout=layout, fused_steps=n_steps)
            launches[lk]()

        self.layout = autotune.best_config(sig, cands, go)
        for x, s in zip(self.x, saved):
            x.copy_(s)
        return self.layout
