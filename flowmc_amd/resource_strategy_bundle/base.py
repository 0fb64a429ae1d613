"""ResourceStrategyBundle — src/flowMC/resource_strategy_bundle/base.py:7-17."""

from abc import ABC


class ResourceStrategyBundle(ABC):
    resources: dict
    strategies: dict
    strategy_order: list
