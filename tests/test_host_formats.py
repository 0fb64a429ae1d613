"""On-disk formats of Buffer and State (src/flowMC/resource/buffers.py:52-64,
states.py:53-63); CPU only."""

import numpy as np
import pytest
import torch

from flowmc_amd.resource.buffers import Buffer
from flowmc_amd.resource.states import State


def test_buffer_file_has_the_reference_layout(tmp_path):
    b = Buffer("positions", (3, 4, 2), 1, device="cpu")
    b.update_buffer(torch.arange(12.0).reshape(3, 2, 2), 1)
    b.save_resource(str(tmp_path) + "/run_")
    f = np.load(str(tmp_path) + "/run_positions.npz", allow_pickle=False)
    assert sorted(f.files) == ["data", "name"]
    assert str(f["name"]) == "positions"
    ref = np.full((3, 4, 2), -np.inf, dtype=np.float32)
    ref[:, 1:3] = np.arange(12.0).reshape(3, 2, 2)
    assert np.array_equal(f["data"], ref)


def test_buffer_loads_a_reference_style_file(tmp_path):
    # np.savez(path + name, name=name, data=data), as buffers.py:52-57 writes it
    data = np.random.default_rng(0).normal(size=(5, 7)).astype(np.float32)
    np.savez(str(tmp_path) + "/x_losses", name="losses", data=data)
    loader = Buffer("other", (1,), 1, device="cpu")
    b = loader.load_resource(str(tmp_path) + "/x_losses.npz")
    assert b.name == "losses" and b.shape == (5, 7)
    assert b.cursor == 0 and b.cursor_dim == 0  # Buffer(data["name"], buffer.shape): defaults
    assert torch.equal(b.data, torch.as_tensor(data))


def test_buffer_refuses_pickled_files(tmp_path):
    np.savez(str(tmp_path) + "/bad", name="bad", data=np.array([{"a": 1}], dtype=object))
    with pytest.raises(ValueError):
        Buffer("x", (1,), device="cpu").load_resource(str(tmp_path) + "/bad.npz")


def test_state_round_trip_without_pickle(tmp_path):
    s = State({"target_positions": "positions_training", "training": True, "n": 3, "w": 0.5}, name="sampler_state")
    s.save_resource(str(tmp_path) + "/")
    f = np.load(str(tmp_path) + "/sampler_state.npz", allow_pickle=False)  # no object arrays
    assert str(f["name"]) == "sampler_state"
    t = s.load_resource(str(tmp_path) + "/sampler_state.npz")
    assert t.name == "sampler_state" and t.data == s.data
    assert isinstance(t.data["training"], bool) and isinstance(t.data["n"], int)
