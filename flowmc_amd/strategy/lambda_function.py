"""Lambda — src/flowMC/strategy/lambda_function.py:7-37 (return value of the function is ignored)."""

from typing import Callable

from .base import Strategy


class Lambda(Strategy):
    def __init__(self, lambda_function: Callable):
        self.lambda_function = lambda_function

    def __call__(self, rng_key, resources, initial_position, data):
        self.lambda_function(rng_key, resources, initial_position, data)
        return rng_key, resources, initial_position
