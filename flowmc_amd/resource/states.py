"""State resource — src/flowMC/resource/states.py:8-63.

A name -> (int | bool | str) dict that routes strategies to buffers.
"""

from __future__ import annotations

import numpy as np

from .base import Resource


class State(Resource):
    name: str
    data: dict

    def __repr__(self):
        return "State " + self.name + " with shape " + str(len(self.data))

    def __init__(self, data: dict, name: str = "State"):
        self.name = name
        self.data = data

    def update(self, key: list, value: list):
        """In-place update of several entries (states.py:37-48)."""
        for k, v in zip(key, value):
            self.data[k] = v
            print(f"Updated state {k} to {v}")

    def print_parameters(self):
        print(f"State: {self.name} with shape {len(self.data)} and data {self.data}")

    def save_resource(self, path: str):
        # states.py:53-58 stores the dict itself, which numpy can only keep as a
        # pickled object array (and states.py:60-63 then cannot load it without
        # allow_pickle); here the dict is stored as parallel string arrays so it
        # loads without pickle
        keys = np.array(list(self.data.keys()), dtype=str)
        vals = np.array([repr(v) for v in self.data.values()], dtype=str)
        np.savez(path + self.name, name=self.name, keys=keys, values=vals)

    def load_resource(self, path: str) -> "State":
        import ast

        f = np.load(path, allow_pickle=False)
        data = {str(k): ast.literal_eval(str(v)) for k, v in zip(f["keys"], f["values"])}
        return State(data, str(f["name"]))
