"""Loads the committed fixtures of tests/golden (written by make_golden.py)."""
import os

import numpy as np

from bhgpu import layout

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["two_body", "shared_ic_20", "box2000", "box6000", "plummer5000"]


def load(name):
    z = np.load(os.path.join(HERE, f"{name}.npz"))
    d = {k: z[k] for k in z.files}
    d["particles"] = np.ascontiguousarray(d["particles"]).view(layout.PARTICLE).reshape(-1)
    d["nodes"] = np.ascontiguousarray(d["nodes"]).view(layout.WALK_NODE).reshape(-1)
    ml, gs, ts = (int(v) for v in d["params"])
    d["max_leaf"], d["group_size"], d["target_size"] = ml, gs, ts
    theta, eps2, G = (float(v) for v in d["force"])
    d["fp"] = layout.force_params(theta, eps2, G, True)
    return d


def bytes_equal(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))
