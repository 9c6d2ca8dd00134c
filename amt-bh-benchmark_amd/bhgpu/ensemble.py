"""The mixed-N ensemble: independent systems of uneven size stepped together.

64 systems, N log-uniform in [4096, 1048576] (N_k = round(4096 * 256^u_k),
u_k the first draw of stream k of seed 2024 in ic's generator), Plummer for
even k and Hernquist for odd k, each from seed 1000 + k; theta cycles
0.4, 0.5, 0.7; eps = 0.01 (eps2 = 1e-4); quadrupoles on.
"""
import numpy as np

from . import ic
from .layout import PARTICLE, SYSTEM, force_params, system

THETAS = (0.4, 0.5, 0.7)


def sizes(count=64, n_min=4096, n_max=1 << 20, seed=2024):
    u = ic.uniform(ic.stream_state(seed, np.arange(count, dtype=np.uint64)), 0)
    return np.rint(n_min * (n_max / n_min) ** u).astype(np.int64)


def make(count=64, n_min=4096, n_max=1 << 20, eps=0.01):
    """Particles (systems back to back) and the per-system parameters
    (theta, eps2, kind, seed) of the ensemble."""
    ns = sizes(count, n_min, n_max)
    parts, meta = [], []
    for k, n in enumerate(ns):
        kind = "plummer" if k % 2 == 0 else "hernquist"
        gen = ic.plummer if kind == "plummer" else ic.hernquist
        parts.append(gen(int(n), 1000 + k))
        meta.append({"n": int(n), "kind": kind, "seed": 1000 + k, "theta": THETAS[k % 3], "eps2": eps * eps})
    return (np.concatenate(parts) if parts else np.zeros(0, PARTICLE)), meta


def table(meta, shapes, kick_dt, drift_dt):
    """bh_system records; shapes(meta entry) -> (max_leaf, group_size,
    target_size[, walk_tile])."""
    rows = []
    for m in meta:
        sh = tuple(shapes(m))
        ml, gs, ts = sh[:3]
        tile = sh[3] if len(sh) > 3 else 0
        rows.append(system(m["n"], force_params(m["theta"], m["eps2"], 1.0, True), ml, gs, ts, kick_dt, drift_dt,
                           walk_tile=tile))
    return np.concatenate(rows) if rows else np.zeros(0, SYSTEM)
