"""flowmc_amd — MI355X (gfx950) native implementation of flowMC's sampling hot path.

Module layout mirrors the reference package (``flowMC.resource``,
``flowMC.strategy``, ``flowMC.resource_strategy_bundle``, ``flowMC.Sampler``);
the arithmetic runs in hand-written HIP kernels behind the C ABI of
``include/flowmc.h`` (``libflowmc_hip.so``, loaded with ctypes).
"""

__version__ = "0.1.0"
