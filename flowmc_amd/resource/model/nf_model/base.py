"""NFModel — src/flowMC/resource/model/nf_model/base.py:14-245.

The training loop keeps the reference's semantics:
  * ``train`` resets data_mean / data_cov from the training data
    (jnp.mean, jnp.cov with ddof=1; base.py:187-188),
  * each epoch permutes the data with a fresh key and drops the trailing
    partial batch, or takes one full-data step when N < batch_size
    (base.py:127-151),
  * the returned model and optimizer state are those of the epoch with the
    lowest last-minibatch loss (base.py:196-200),
  * ``train`` returns a new model object; the caller's model is unchanged.
"""

from __future__ import annotations

from abc import abstractmethod

import torch

from .... import random as frandom
from ...base import Resource


def data_stats(data: torch.Tensor):
    """jnp.mean(data, axis=0) and jnp.cov(data.T) (ddof 1) of train()
    (base.py:187-188) without torch.cov, whose degrees-of-freedom check reads
    a device scalar back (a host sync before every training). The rows are
    split into up to 256 chunks whose d x d products run as one batched GEMM
    and are then summed: a single GEMM over 100k rows ran on a handful of
    workgroups (285 us for d = 64)."""
    n, d = data.shape
    mean = data.mean(0)
    xc = data - mean
    p = max(1, min(256, n // 256))
    c = n // p
    head = xc[: p * c].reshape(p, c, d)
    cov = torch.bmm(head.transpose(1, 2).contiguous(), head).sum(0)
    if p * c < n:
        tail = xc[p * c:]
        cov = cov + tail.T.contiguous() @ tail
    return mean, cov / max(n - 1, 1)


class NFModel(Resource):
    @property
    def n_features(self):
        return self._n_features

    @property
    @abstractmethod
    def data_mean(self):
        raise NotImplementedError

    @property
    @abstractmethod
    def data_cov(self):
        raise NotImplementedError

    def __call__(self, x):
        return self.forward(x)

    @abstractmethod
    def log_prob(self, x):
        raise NotImplementedError

    @abstractmethod
    def sample(self, rng_key, n_samples: int):
        raise NotImplementedError

    @abstractmethod
    def forward(self, x, key=None):
        raise NotImplementedError

    @abstractmethod
    def inverse(self, x):
        raise NotImplementedError

    @abstractmethod
    def clone(self) -> "NFModel":
        raise NotImplementedError

    @abstractmethod
    def set_data_stats(self, mean: torch.Tensor, cov: torch.Tensor):
        raise NotImplementedError

    def loss_fn(self, x):
        """-mean(log_prob(x)) (base.py:98-100)."""
        return -self.log_prob(x).mean()

    @abstractmethod
    def train_step(self, x, optim, state):
        """One optimiser step on batch x; returns (loss tensor, model, state)."""

    def train_epoch(self, rng, optim, state, data, batch_size):
        """base.py:127-151 (host loop over shuffled minibatches)."""
        value = torch.tensor(1e9)
        model = self
        n = data.shape[0]
        steps = n // batch_size
        if steps > 0:
            perms = frandom.permutation(rng, n, device=data.device)[: steps * batch_size].reshape(steps, batch_size)
            batches = data[perms]  # the epoch's minibatches in one gather
            # only the last step's loss is returned: the steps before it skip
            # the loss reduction where the model offers that
            quiet = getattr(model, "_train_step_no_loss", None)
            for i in range(steps):
                if quiet is not None and i + 1 < steps:
                    model, state = quiet(batches[i], optim, state)
                    quiet = getattr(model, "_train_step_no_loss", None)
                else:
                    value, model, state = model.train_step(batches[i], optim, state)
        else:
            value, model, state = model.train_step(data, optim, state)
        return value, model, state

    def train(self, rng, data, optim, state, num_epochs: int, batch_size: int, verbose: bool = True):
        """base.py:153-210. Returns (rng, best_model, best_state, loss_values[num_epochs])."""
        data = torch.as_tensor(data, dtype=torch.float32)
        if not data.is_cuda:
            data = data.cuda()
        state0 = state
        if hasattr(self, "_training_copy"):  # working buffers reused from call to call
            model, state = self._training_copy(state)
        else:
            model, state = self.clone(), state.clone()
        mean, cov = data_stats(data)
        model.set_data_stats(mean, cov.reshape(self.n_features, self.n_features))
        it = range(num_epochs)
        if verbose:
            try:
                from tqdm import trange

                it = trange(num_epochs, desc="Training NF", miniters=max(1, int(num_epochs / 10)))
            except ImportError:  # pragma: no cover
                pass
        # The best epoch (the first of lowest last-minibatch loss, base.py:196-200)
        # is tracked on the device: its parameters and Adam moments are kept by
        # selects, and the losses are read back once after the loop instead of
        # a host sync per epoch (verbose reads them as the reference prints).
        dev = data.device
        losses = torch.empty(num_epochs, dtype=torch.float32, device=dev)
        best_loss = torch.full((), 1e9, dtype=torch.float32, device=dev)
        best_epoch = torch.full((), -1, dtype=torch.float32, device=dev)
        best = None
        counts = []
        for epoch in it:
            rng, input_rng = frandom.split(rng)
            value, model, state = model.train_epoch(input_rng, optim, state, data, batch_size)
            losses[epoch] = torch.as_tensor(value, dtype=torch.float32, device=dev)
            counts.append(state.count)
            better = losses[epoch] < best_loss
            if best is None:
                best = [model.theta.clone(), state.mu.clone(), state.nu.clone()]
            for dst, src in zip(best, (model.theta, state.mu, state.nu)):
                torch.where(better, src, dst, out=dst)
            best_loss = torch.where(better, losses[epoch], best_loss)
            best_epoch = torch.where(better, torch.full_like(best_epoch, epoch), best_epoch)
            if verbose and hasattr(it, "set_description"):
                if num_epochs > 10 and epoch % int(num_epochs / 10) == 0:
                    it.set_description(f"Training NF, current loss: {float(losses[epoch]):.3f}")
        back = torch.cat([losses, best_epoch.reshape(1)]).cpu()
        loss_values, be = back[:num_epochs].clone(), int(back[num_epochs])
        if be < 0:
            best_model, best_state = self, state0
        else:
            best_model = model.clone(best[0])
            best_state = type(state)(best[1], best[2], counts[be])
        return rng, best_model, best_state, loss_values

    def save_resource(self, path):
        return self.save_model(path)

    def load_resource(self, path):
        return self.load_model(path)
