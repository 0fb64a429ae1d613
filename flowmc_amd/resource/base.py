"""Resource ABC — src/flowMC/resource/base.py:6-38."""

from abc import ABC, abstractmethod


class Resource(ABC):
    """Base class for resources (kernels, models, buffers, states...)."""

    @abstractmethod
    def __init__(self):
        raise NotImplementedError

    @abstractmethod
    def print_parameters(self):
        """Print the tunable parameters of the resource."""
        raise NotImplementedError

    @abstractmethod
    def save_resource(self, path: str):
        raise NotImplementedError

    @abstractmethod
    def load_resource(self, path: str):
        raise NotImplementedError
