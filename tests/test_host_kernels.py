"""Host-side behaviour of the proposal resources that needs no GPU: repr and
print_parameters strings the reference's tests assert (test/unit/test_kernels.py)."""

from flowmc_amd.resource.kernel.Gaussian_random_walk import GaussianRandomWalk
from flowmc_amd.resource.kernel.MALA import MALA


def test_grw_repr_and_parameters(capsys):
    k = GaussianRandomWalk(step_size=1)
    assert repr(k) == "Gaussian Random Walk with step size 1"
    k.print_parameters()
    assert capsys.readouterr().out == "Gaussian Random Walk parameters:\nstep_size: 1\n"


def test_mala_repr_and_parameters(capsys):
    k = MALA(step_size=1)
    assert repr(k) == "MALA with step size 1"
    k.print_parameters()
    assert capsys.readouterr().out == "MALA parameters:\nstep_size: 1\n"
