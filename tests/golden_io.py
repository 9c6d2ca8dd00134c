"""Loads the committed fixtures of tests/golden (written by make_golden.py)."""
import hashlib
import json
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


DIGEST_NAMES = ["plummer131k"]


def load_digest(name):
    """A digest fixture: the seeded input's description and SHA-256 digests
    of the reference's outputs for it."""
    return json.load(open(os.path.join(HERE, f"{name}.json")))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()


def digest_input(d):
    """Regenerates a digest fixture's input (ic.plummer(n, seed)) and checks it."""
    from bhgpu import ic

    assert d["input"].startswith("ic.plummer(")
    n, seed = (int(v) for v in d["input"][len("ic.plummer("):-1].split(","))
    p = ic.plummer(n, seed)
    assert sha(p) == d["input_sha256"], "the seeded generator no longer gives the fixture's input"
    return p
