"""The committed golden vectors are what the pinned oracle produces (CPU)."""

import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["spline", "flow", "mala", "hmc", "nf_mh", "adamw"])
def test_fixture_matches_oracle(name):
    import importlib.util

    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    fresh = mg.BUILDERS[name]()
    stored = np.load(os.path.join(HERE, f"{name}.npz"))
    assert set(stored.files) == set(fresh)
    for k in stored.files:
        np.testing.assert_allclose(stored[k], np.asarray(fresh[k]), rtol=1e-10, atol=1e-10, err_msg=f"{name}.{k}")
