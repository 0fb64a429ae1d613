"""Strategy ABC — src/flowMC/strategy/base.py:8-31."""

from abc import ABC, abstractmethod


class Strategy(ABC):
    """A block that transforms (rng_key, resources, positions)."""

    @abstractmethod
    def __init__(self):
        raise NotImplementedError

    @abstractmethod
    def __call__(self, rng_key, resources: dict, initial_position, data: dict):
        """Returns (rng_key, resources, last_position)."""
        raise NotImplementedError
