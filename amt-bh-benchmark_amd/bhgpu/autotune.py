"""Autotuner for the walk shapes: leaf capacity, group size, target size and
walk tile.

Leaf capacity, group size and target size change the tree and the targets'
boxes, so they change both the work and the interaction lists (hence the
accuracy); the walk tile (particles per walk unit) changes only how targets
share walks, i.e. the time. For each size class and opening angle the tuner
measures, on a Plummer and a Hernquist system of that size (eps = 0.01):

- the rms relative acceleration error of every (leaf, group, target) shape
  against the FP64 direct sum (seeded subset, worst of two Plummer and two
  Hernquist systems);
- the step time (median of a few steps, summed over a Plummer and a
  Hernquist system) of every walk tile for each shape within the target.

The error is a hard constraint with a margin: the fastest (shape, tile)
whose error is within MARGIN times the bound wins, so that other seeds and
nearby sizes, whose errors scatter by a fifth or more, stay within the
bound. When no shape reaches the target the most accurate shape is kept;
the class records whether that shape is within the bound itself
(within_bound) and the error floor of all shapes.

The table ((size class, theta) -> shapes and the measurements behind the
choice) is written as JSON next to this module; `shapes_for(n, theta)`
reads it (nearest size class in log N, nearest theta) and falls back to
DEFAULT when no table was produced.

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
MAX_TILE = 128
DEFAULT = (16, 256, 32, MAX_TILE)  # (max_leaf, group_size, target_size, walk_tile)
LEAVES = (8, 16, 32, 64)
GROUPS = (64, 128, 256)
TARGETS = (16, 32, 64, 96, 128, 256)
TILES = (32, 64, 128)
MARGIN = 0.8

_cache = {}


def load_table(path=TABLE):
    if not os.path.exists(path):
        return None
    key = (path, os.path.getmtime(path))
    if key not in _cache:
        _cache[key] = json.load(open(path))
    return _cache[key]


def class_for(n, theta=0.5, path=TABLE):
    """The table's entry for a system of n particles (nearest size class in
    log N, then nearest theta), or None without a table."""
    t = load_table(path)
    if not t or not t.get("classes"):
        return None
    return min(t["classes"], key=lambda c: (abs(np.log(max(n, 1)) - np.log(c["n"])), abs(theta - c.get("theta", theta))))


def shapes_for(n, theta=0.5, path=TABLE):
    """(max_leaf, group_size, target_size, walk_tile) for a system of n
    particles; tables without a tile give the largest tile."""
    best = class_for(n, theta, path)
    if best is None:
        return DEFAULT
    sh = tuple(best["shapes"])
    return sh if len(sh) == 4 else sh + (MAX_TILE,)


def _accuracy(ctx, p, fp, shapes, subset=256):
    ml, gs, ts = shapes[:3]
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
    ml, gs, ts, tile = shapes
    ctx.set_walk_tile(tile)
    try:
        ctx.state_upload(p)
        ctx.state_step(fp, ml, gs, ts, 0.0, 0.0)  # warm-up, grows buffers
        ts_ms = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ctx.state_step(fp, ml, gs, ts, 0.0, 0.0)
            ts_ms.append(1e3 * (time.perf_counter() - t0))
    finally:
        ctx.set_walk_tile(MAX_TILE)
    return float(np.median(ts_ms))


def tune(ctx, sizes, thetas=(0.4, 0.5, 0.7), eps=0.01, bound=1e-3, margin=MARGIN, log=print):
    classes = []
    for n in sizes:
        systems = [ic.plummer(n, 31), ic.hernquist(n, 32), ic.plummer(n, 33), ic.hernquist(n, 34)]
        for theta in thetas:
            fp = layout.force_params(theta, eps * eps, 1.0, True)
            shapes = [(ml, gs, ts) for ml in LEAVES for gs in GROUPS for ts in TARGETS if ts <= gs]
            errs = {sh: max(_accuracy(ctx, p, fp, sh) for p in systems) for sh in shapes}
            ok = [sh for sh in shapes if errs[sh] <= margin * bound]
            timed = ok if ok else [min(shapes, key=lambda sh: errs[sh])]
            rows = []
            for sh in timed:
                for tile in TILES:
                    ms = sum(_time(ctx, p, fp, sh + (tile,)) for p in systems[:2])
                    rows.append({"shapes": list(sh) + [tile], "ms": ms, "rms_rel": errs[sh]})
            best = min(rows, key=lambda r: r["ms"])
            dflt_err = errs.get(DEFAULT[:3])
            log(f"n={n} theta={theta}: best {best['shapes']} {best['ms']:.3f} ms (rms {best['rms_rel']:.2e}, "
                f"{len(ok)} of {len(shapes)} shapes within {margin:g} x {bound:g}; "
                f"error floor {min(errs.values()):.2e})")
            classes.append({"n": int(n), "theta": theta, "shapes": best["shapes"], "ms": best["ms"],
                            "rms_rel": best["rms_rel"], "within_target": bool(ok),
                            "within_bound": best["rms_rel"] <= bound,
                            "rms_floor": min(errs.values()), "default_rms_rel": dflt_err,
                            "shape_errors": [{"shapes": list(sh), "rms_rel": e} for sh, e in errs.items()],
                            "timings": rows})
    return {"eps": eps, "bound": bound, "margin": margin,
            "measure": "rms: worst of two Plummer and two Hernquist systems (FP64 direct sum, 256 particles "
                       "each); time: median of 3 steps, summed over a Plummer and a Hernquist system, for "
                       "every walk tile of every shape within margin x bound",
            "classes": classes}


def dump(table, f):
    """The table as JSON with one size class per line."""
    head = {k: v for k, v in table.items() if k != "classes"}
    f.write(json.dumps(head)[:-1] + ', "classes": [\n')
    f.write(",\n".join(json.dumps(c, separators=(",", ":")) for c in table["classes"]))
    f.write("\n]}\n")


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
        dump(table, f)
    print(json.dumps({"classes": [{k: c[k] for k in ("n", "theta", "shapes", "ms", "rms_rel", "within_bound", "rms_floor")}
                                  for c in table["classes"]]}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
