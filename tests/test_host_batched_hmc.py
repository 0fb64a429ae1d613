"""Host-side validation of the batched uneven-dimension HMC tables: the kernel
reads its slot table from device memory, so every shape, row range and
target descriptor is checked on the host before any launch. CPU only: the
checks raise before the first device pointer is taken."""

import pytest
import torch

from flowmc_amd import targets as TG
from flowmc_amd.resource.kernel.batched_HMC import BatchedHMC


def _bhmc(dims=(8, 20), chains=(16, 8)):
    b = object.__new__(BatchedHMC)  # no GPU here: skip __init__'s device check
    b.targets = [TG.Funnel(d, device="cpu") for d in dims]
    b.n_chains = list(chains)
    b.dims = list(dims)
    b.step_size, b.n_leapfrog = 0.05, 10
    b.device = torch.device("cpu")
    b.metrics = [None] * len(dims)
    b.x = [torch.zeros(c, d) for c, d in zip(chains, dims)]
    b.logp = [torch.zeros(c) for c in chains]
    b.layout = None
    return b


def _outs(b, T=10, cursor=0):
    return [(torch.zeros(c, T, d), torch.zeros(c, T), torch.zeros(c, T), cursor) for c, d in zip(b.n_chains, b.dims)]


def _tables(b, outputs, noise=None, log_u=None, n_steps=4, thinning=1, step0=0):
    lay = b.default_layout()
    return b._tables(lay, (0, 1), n_steps, outputs, noise, log_u, thinning, step0, n_steps)


def test_rows_past_the_buffer_end_are_refused():
    b = _bhmc()
    with pytest.raises(ValueError, match="do not fit a buffer of length 10"):
        _tables(b, _outs(b, T=10, cursor=8), n_steps=4)
    with pytest.raises(ValueError, match="do not fit"):
        _tables(b, _outs(b, T=3), n_steps=10, thinning=3)  # rows 0..3


def test_buffer_length_from_any_given_output_and_agreement():
    b = _bhmc()
    outs = _outs(b)
    # positions absent: the length comes from the log_prob / acceptance buffers
    outs = [(None, lp[:, :2], acc[:, :2], 0) for _, lp, acc, _ in outs]
    with pytest.raises(ValueError, match="do not fit a buffer of length 2"):
        _tables(b, outs, n_steps=4)
    outs = _outs(b)
    outs[1] = (outs[1][0], outs[1][1][:, :5], outs[1][2], 0)
    with pytest.raises(ValueError, match="disagree on their length"):
        _tables(b, outs)


def test_buffer_shapes_noise_metric_and_state():
    b = _bhmc()
    outs = _outs(b)
    outs[0] = (torch.zeros(16, 10, 7), outs[0][1], outs[0][2], 0)
    with pytest.raises(ValueError, match="slot 0: position buffer must be"):
        _tables(b, outs)
    noise = [torch.zeros(4, c, d) for c, d in zip(b.n_chains, b.dims)]
    log_u = [torch.zeros(4, c) for c in b.n_chains]
    noise[1] = torch.zeros(3, 8, 20)
    with pytest.raises(ValueError, match="slot 1: noise must be"):
        _tables(b, _outs(b), noise, log_u, n_steps=4)
    noise[1] = torch.zeros(4, 8, 20)
    log_u[0] = torch.zeros(4, 15)
    with pytest.raises(ValueError, match="slot 0: log_u must be"):
        _tables(b, _outs(b), noise, log_u, n_steps=4)
    with pytest.raises(ValueError, match="injected together"):
        _tables(b, _outs(b), noise, None, n_steps=4)
    b.metrics[1] = torch.ones(19)
    with pytest.raises(ValueError, match="slot 1: metric must be"):
        _tables(b, _outs(b))
    b.metrics[1] = None
    b.x[0] = torch.zeros(16, 9)
    with pytest.raises(ValueError, match="slot 0: state must be"):
        _tables(b, _outs(b))


def test_target_descriptor_is_checked_on_the_host():
    from flowmc_amd import _native as N

    b = _bhmc()
    b.targets[1] = TG.GaussianDiag(torch.zeros(20), torch.ones(20), device="cpu")
    b.targets[1]._n_params = 39  # needs 2 d = 40 floats
    with pytest.raises(N.FlowMCNativeError, match="needs 40 params"):
        _tables(b, _outs(b))
    b.targets[1] = TG.Funnel(21, device="cpu")
    with pytest.raises(ValueError, match="target n_dim 21 != 20"):
        _tables(b, _outs(b))


def test_valid_tables_reach_the_device_pointer_stage():
    # everything above passes: the first device-pointer check is what stops a CPU run
    b = _bhmc()
    with pytest.raises(ValueError, match="must live on the HIP device"):
        _tables(b, _outs(b))
