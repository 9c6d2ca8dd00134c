"""Autotuner for the walk shapes (leaf capacity, group size, target size).

The three shapes change the tree and the targets' boxes, so they change
both the work and the interaction lists (hence the accuracy). For each size
class and opening angle the tuner times every candidate on the GPU (median
of a few steps of a Plummer and a Hernquist system of that size,
eps = 0.01) and keeps the fastest whose rms relative acceleration error
against the FP64 direct sum (seeded subset, worse of the two systems) is
within the bound; when none is, the most accurate one.

The table ((size class, theta) -> shapes and the measurements behind the
choice) is written as JSON next to this module; `shapes_for(n, theta)`
reads it (nearest size class in log N, nearest theta) and falls back to
(16, 256, 32) when no table was produced.

    python -m bhgpu.autotune [--out PATH] [--sizes 4096,32768,...]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import ic, layout
from . import lib as bl

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
DEFAULT = (16, 256, 32)
LEAVES = (8, 16, 32, 64)
GROUPS = (64, 128, 256)
TARGETS = (16, 32, 64)

_cache = {}


def load_table(path=TABLE):
    if not os.path.exists(path):
        return None
    key = (path, os.path.getmtime(path))
    if key not in _cache:
        _cache[key] = json.load(open(path))
    return _cache[key]


def shapes_for(n, theta=0.5, path=TABLE):
    """(max_leaf, group_size, target_size) for a system of n particles."""
    t = load_table(path)
    if not t or not t.get("classes"):
        return DEFAULT
    best = min(t["classes"], key=lambda c: (abs(np.log(max(n, 1)) - np.log(c["n"])), abs(theta - c.get("theta", theta))))
    return tuple(best["shapes"])


def _accuracy(ctx, p, fp, shapes, subset=256):
    ml, gs, ts = shapes
    ctx.build_tree(p, ml, gs)
    acc, _, _, _ = ctx.accel(fp, ts)
    perm = ctx.export_perm()
    where = np.empty(len(p), np.int64)
    where[perm] = np.arange(len(p))
    rng = np.random.default_rng(99)
    idx = np.sort(rng.choice(len(p), size=min(subset, len(p)), replace=False)).astype(np.uint32)
    ref = ctx.direct_sum(p, idx, float(fp["eps2"][0]), 1.0)
    rel = np.linalg.norm(acc[where[idx]] - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return float(np.sqrt(np.mean(rel ** 2)))


def _time(ctx, p, fp, shapes, reps=3):
    ml, gs, ts = shapes
    ctx.state_upload(p)
    ctx.state_step(fp, ml, gs, ts, 0.0, 0.0)  # warm-up, grows buffers
    ts_ms = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.state_step(fp, ml, gs, ts, 0.0, 0.0)
        ts_ms.append(1e3 * (time.perf_counter() - t0))
    return float(np.median(ts_ms))


def tune(ctx, sizes, thetas=(0.4, 0.5, 0.7), eps=0.01, bound=1e-3, log=print):
    classes = []
    for n in sizes:
        systems = [ic.plummer(n, 31), ic.hernquist(n, 32)]
        for theta in thetas:
            fp = layout.force_params(theta, eps * eps, 1.0, True)
            rows = []
            for ml in LEAVES:
                for gs in GROUPS:
                    for ts in TARGETS:
                        if ts > gs:
                            continue
                        sh = (ml, gs, ts)
                        ms = sum(_time(ctx, p, fp, sh) for p in systems)
                        err = max(_accuracy(ctx, p, fp, sh) for p in systems)
                        rows.append({"shapes": sh, "ms": ms, "rms_rel": err})
            ok = [r for r in rows if r["rms_rel"] <= bound]
            best = min(ok, key=lambda r: r["ms"]) if ok else min(rows, key=lambda r: r["rms_rel"])
            dflt = next((r for r in rows if tuple(r["shapes"]) == DEFAULT), None)
            log(f"n={n} theta={theta}: best {best['shapes']} {best['ms']:.3f} ms (rms {best['rms_rel']:.2e}, "
                f"{len(ok)} of {len(rows)} within {bound:g}); default {dflt['ms']:.3f} ms (rms {dflt['rms_rel']:.2e})")
            classes.append({"n": int(n), "theta": theta, "shapes": list(best["shapes"]), "ms": best["ms"],
                            "rms_rel": best["rms_rel"], "within_bound": bool(ok), "default_ms": dflt["ms"],
                            "default_rms_rel": dflt["rms_rel"], "candidates": rows})
    return {"eps": eps, "bound": bound, "measure": "median of 3 steps each of a Plummer and a Hernquist system",
            "classes": classes}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--out", default=TABLE)
    ap.add_argument("--sizes", default="4096,16384,65536,262144,1048576")
    ap.add_argument("--thetas", default="0.4,0.5,0.7")
    args = ap.parse_args(argv)
    ctx = bl.Context(0)
    table = tune(ctx, [int(s) for s in args.sizes.split(",")], thetas=[float(t) for t in args.thetas.split(",")],
                 log=lambda *a: print(*a, file=sys.stderr, flush=True))
    ctx.close()
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print(json.dumps({"classes": [{k: c[k] for k in ("n", "theta", "shapes", "ms", "rms_rel", "default_ms")}
                                  for c in table["classes"]]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
