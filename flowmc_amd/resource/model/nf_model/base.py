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
            for perm in perms:
                value, model, state = model.train_step(data[perm], optim, state)
        else:
            value, model, state = model.train_step(data, optim, state)
        return value, model, state

    def train(self, rng, data, optim, state, num_epochs: int, batch_size: int, verbose: bool = True):
        """base.py:153-210. Returns (rng, best_model, best_state, loss_values[num_epochs])."""
        data = torch.as_tensor(data, dtype=torch.float32)
        if not data.is_cuda:
            data = data.cuda()
        loss_values = torch.zeros(num_epochs, dtype=torch.float32)
        best_model, best_state, best_loss = self, state, 1e9
        model = self.clone()
        state = state.clone()
        model.set_data_stats(data.mean(0), torch.cov(data.T).reshape(self.n_features, self.n_features))
        it = range(num_epochs)
        if verbose:
            try:
                from tqdm import trange

                it = trange(num_epochs, desc="Training NF", miniters=max(1, int(num_epochs / 10)))
            except ImportError:  # pragma: no cover
                pass
        for epoch in it:
            rng, input_rng = frandom.split(rng)
            value, model, state = model.train_epoch(input_rng, optim, state, data, batch_size)
            v = float(value)
            loss_values[epoch] = v
            if v < best_loss:
                best_model, best_state, best_loss = model.clone(), state.clone(), v
            if verbose and hasattr(it, "set_description"):
                if num_epochs > 10 and epoch % int(num_epochs / 10) == 0:
                    it.set_description(f"Training NF, current loss: {v:.3f}")
        return rng, best_model, best_state, loss_values

    def save_resource(self, path):
        return self.save_model(path)

    def load_resource(self, path):
        return self.load_model(path)
