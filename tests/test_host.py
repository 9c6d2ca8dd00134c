"""Host-side logic: the ensemble definition, the autotuner's table lookup,
the server's footprint model and bench.py's one-GPU contract (no GPU)."""
import json
import os
import subprocess
import sys

import numpy as np

from bhgpu import autotune, ensemble
from bhgpu import lib as bl

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_ensemble_sizes_log_uniform():
    ns = ensemble.sizes()
    assert len(ns) == 64 and ns.min() >= 4096 and ns.max() <= 1 << 20
    assert np.array_equal(ns, ensemble.sizes())  # seeded
    u = np.log(ns / 4096) / np.log(256)  # uniform in [0, 1)
    assert abs(u.mean() - 0.5) < 0.1 and u.min() < 0.1 and u.max() > 0.9


def test_ensemble_small_make_and_table():
    p, meta = ensemble.make(count=6, n_min=100, n_max=400)
    assert len(p) == sum(m["n"] for m in meta)
    assert [m["kind"] for m in meta] == ["plummer", "hernquist"] * 3
    assert [m["theta"] for m in meta] == [0.4, 0.5, 0.7] * 2
    t = ensemble.table(meta, lambda m: (8, 64, 16), 1e-3, 2e-3)
    assert list(t["n"]) == [m["n"] for m in meta] and np.all(t["max_leaf"] == 8) and np.all(t["kick_dt"] == 1e-3)
    assert np.allclose(t["fp"]["eps2"], 1e-4)


def test_autotune_lookup(tmp_path):
    path = tmp_path / "t.json"
    assert autotune.shapes_for(1000, 0.5, str(path)) == autotune.DEFAULT
    json.dump({"classes": [{"n": 4096, "theta": 0.5, "shapes": [8, 64, 16, 32]},
                           {"n": 4096, "theta": 0.7, "shapes": [16, 64, 16, 64]},
                           {"n": 1 << 20, "theta": 0.5, "shapes": [32, 256, 32]}]}, open(path, "w"))
    assert autotune.shapes_for(5000, 0.5, str(path)) == (8, 64, 16, 32)
    assert autotune.shapes_for(5000, 0.7, str(path)) == (16, 64, 16, 64)
    # a table without the tile axis gives the largest tile
    assert autotune.shapes_for(500000, 0.4, str(path)) == (32, 256, 32, autotune.MAX_TILE)


def test_server_footprint_model():
    lib = bl.load()
    f16 = lib.bh_server_footprint(bl.ctypes.byref(bl.request(1 << 20, max_leaf=16)))
    f8 = lib.bh_server_footprint(bl.ctypes.byref(bl.request(1 << 20, max_leaf=8)))
    assert 350 * (1 << 20) < f16 < f8 < 1000 * (1 << 20)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
