"""RQSpline_MALA_PT_Bundle — src/flowMC/resource_strategy_bundle/RQSpline_MALA_PT.py:24-391.

The RQSpline_MALA bundle plus parallel tempering: a TemperedPDF resource
(the target as log-likelihood, ``logprior`` as log prior), the tempered walker
buffer [n_chains, n_temperatures - 1, n_dims], the temperature ladder
(``linspace(1, max_temperature, n_temperatures)``), and a ParallelTempering
strategy that runs first in every training and production loop, after an
initial strategy that copies the initial positions into every tempered slot.

``logprior=None`` (the default) is the zero prior of the reference's default
``lambda x, _: 0.0``; it keeps the tempered steps of built-in targets in the
fused chain kernel. Extra keyword arguments as in RQSpline_MALA_Bundle.
"""

from __future__ import annotations

from typing import Callable, Optional

import torch

from ..resource.buffers import Buffer
from ..resource.logPDF import TemperedPDF
from ..strategy.lambda_function import Lambda
from ..strategy.parallel_tempering import ParallelTempering
from .RQSpline_MALA import RQSpline_MALA_Bundle


class RQSpline_MALA_PT_Bundle(RQSpline_MALA_Bundle):
    def __repr__(self):
        return "RQSpline MALA PT Bundle"

    def __init__(self, rng_key, n_chains: int, n_dims: int, logpdf: Callable, n_local_steps: int,
                 n_global_steps: int, n_training_loops: int, n_production_loops: int, n_epochs: int,
                 mala_step_size: float = 1e-1, chain_batch_size: int = 0, rq_spline_hidden_units: list = [32, 32],
                 rq_spline_n_bins: int = 8, rq_spline_n_layers: int = 4, learning_rate: float = 1e-3,
                 batch_size: int = 10000, n_max_examples: int = 10000, local_thinning: int = 1,
                 global_thinning: int = 1, n_NFproposal_batch_size: int = 10000, history_window: int = 100,
                 n_temperatures: int = 5, max_temperature: float = 5.0, n_tempered_steps: int = -1,
                 logprior: Optional[Callable] = None, verbose: bool = False, precision: str = "bfloat16",
                 fused_steps: int = 64, exact_inverse: bool = False):
        super().__init__(rng_key, n_chains, n_dims, logpdf, n_local_steps, n_global_steps, n_training_loops,
                         n_production_loops, n_epochs, mala_step_size=mala_step_size,
                         chain_batch_size=chain_batch_size, rq_spline_hidden_units=rq_spline_hidden_units,
                         rq_spline_n_bins=rq_spline_n_bins, rq_spline_n_layers=rq_spline_n_layers,
                         learning_rate=learning_rate, batch_size=batch_size, n_max_examples=n_max_examples,
                         local_thinning=local_thinning, global_thinning=global_thinning,
                         n_NFproposal_batch_size=n_NFproposal_batch_size, verbose=verbose, precision=precision,
                         fused_steps=fused_steps, exact_inverse=exact_inverse, history_window=history_window)
        tempered_logpdf = TemperedPDF(self.resources["logpdf"], logprior, n_dims=n_dims, n_temps=n_temperatures)
        tempered_positions = Buffer("tempered_positions", (n_chains, n_temperatures - 1, n_dims), 2)
        temperatures = Buffer("temperature", (n_temperatures,), 0)
        ladder = torch.linspace(1.0, float(max_temperature), n_temperatures, dtype=torch.float64)
        temperatures.update_buffer(ladder.to(torch.float32))
        self.resources.update({
            "tempered_logpdf": tempered_logpdf,
            "tempered_positions": tempered_positions,
            "temperatures": temperatures,
        })

        if n_tempered_steps <= 0:
            print("n_tempered_steps value is not valid. Setting to n_local_steps")
            n_tempered_steps = n_local_steps
        parallel_tempering_strat = ParallelTempering(
            n_steps=n_tempered_steps,
            tempered_logpdf_name="tempered_logpdf",
            kernel_name="local_sampler",
            tempered_buffer_names=["tempered_positions", "temperatures"],
            state_name="sampler_state",
            verbose=verbose,
        )

        def initialize_tempered_positions(rng_key, resources, initial_position, data):
            x0 = torch.as_tensor(initial_position, dtype=torch.float32)
            tempered_positions.update_buffer(x0[:, None].expand(-1, n_temperatures - 1, -1))
            return rng_key, resources, initial_position

        self.strategies.update({
            "parallel_tempering": parallel_tempering_strat,
            "initialize_tempered_positions": Lambda(initialize_tempered_positions),
        })

        training_phase = ["parallel_tempering", "local_stepper", "update_global_step", "model_trainer",
                          "update_model", "global_stepper", "update_local_step"]
        production_phase = ["parallel_tempering", "local_stepper", "update_global_step", "global_stepper",
                            "update_local_step"]
        strategy_order = ["initialize_tempered_positions"]
        for _ in range(n_training_loops):
            strategy_order.extend(training_phase)
        strategy_order.append("reset_steppers")
        strategy_order.append("update_state")
        for _ in range(n_production_loops):
            strategy_order.extend(production_phase)
        self.strategy_order = strategy_order
