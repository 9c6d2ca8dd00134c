"""Benchmark: particles/s per step of the Barnes-Hut force path on one MI355X.

Workload (the reference configuration the metric is quoted on): Plummer
sphere N = 16,777,216 (G = M = a = 1, r < 10a, seed 3), eps = 1e-3,
theta = 0.5, leaf capacity 32, group size 256 and target size 32 (the
reference driver's defaults, main.cc:30-33), leapfrog dt = 5e-4, FP32 walk
with compensated accumulation, FP64 tree / multipoles / kick-drift.

A step is one step_particles call (tree build + walk + fused kick/drift) on
particles resident in HBM. The leapfrog's priming half-kick is the first
warm-up step; the K timed steps are full kick+drift steps.

Prints one JSON line. `roofline` is for the walk kernel (the dominant one):
algorithmic FP32 flops per launch (19 per particle interaction, 55 per
quadrupole-cell interaction) over its average duration from HIP events on
the library's stream. `cpu_baseline` times the FP64 oracle (a C++
restatement of the reference's grouped path) on the host's cores over a
bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "amt-bh-benchmark_amd"))

from bhgpu import ic, layout  # noqa: E402
from bhgpu import lib as bl  # noqa: E402

METRIC = "particles/s per step, 16M Plummer θ=0.5 @rms err≤1e-3; p50/p99 batch latency"
PEAK_FP32_TFLOPS = 157.3  # MI355X vector FP32, AMD published
FLOP_PP = 19
FLOP_PC = 55


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rel_errors(test, ref):
    num = np.linalg.norm(test - ref, axis=1)
    den = np.linalg.norm(ref, axis=1)
    rel = num / np.where(den > 0, den, 1.0)
    glob = float(np.sqrt((num ** 2).sum() / (den ** 2).sum()))
    return float(np.sqrt(np.mean(rel ** 2))), float(np.percentile(rel, 99)), glob


def accuracy(ctx, p, fp, args):
    """FP32 tree accelerations vs the FP64 direct sum on a seeded subset."""
    ctx.build_tree(p, args.max_leaf, args.group_size)
    acc, _, _, _ = ctx.accel(fp, args.target_size)
    perm = ctx.export_perm()  # Morton position -> input index
    rng = np.random.default_rng(12345)
    idx = np.sort(rng.choice(len(p), size=min(args.error_subset, len(p)), replace=False)).astype(np.uint32)
    where = np.empty(len(p), np.int64)
    where[perm] = np.arange(len(p))
    ref = ctx.direct_sum(p, idx, fp["eps2"][0], fp["G"][0])
    rms, p99, glob = rel_errors(acc[where[idx]], ref)
    return {"rms_rel": rms, "p99_rel": p99, "global_rel": glob, "subset": int(len(idx)), "reference": "FP64 direct sum"}


def cpu_baseline(p, fp, args):
    sys.path.insert(0, REPO)
    import oracle  # the FP64 restatement, timed as the CPU baseline

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    o = oracle.Oracle()
    build_ms, walk_ms, frac = o.baseline_sample(p, fp, args.max_leaf, args.group_size, args.target_size,
                                                args.cpu_stride, cores)
    step_s = build_ms / 1e3 + (walk_ms / 1e3) / frac
    return {
        "value": len(p) / step_s,
        "unit": "particles/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"FP64 oracle, full tree build of all {len(p)} particles ({build_ms:.0f} ms) + grouped walk of "
                   f"every {args.cpu_stride}th group ({100 * frac:.2f}% of particles, {walk_ms:.0f} ms), walk time "
                   f"scaled to all groups"),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--eps", type=float, default=1e-3)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--max-leaf", type=int, default=32)
    ap.add_argument("--group-size", type=int, default=256)
    ap.add_argument("--target-size", type=int, default=32)
    ap.add_argument("--fp64", action="store_true", help="FP64 walk arithmetic")
    ap.add_argument("--error-subset", type=int, default=4096)
    ap.add_argument("--cpu-stride", type=int, default=128)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (torch) three times, to calibrate HBM byte counters")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch

        a = torch.ones(args.copy_calibration * (1 << 18), dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": a.numel() * 4, "bytes_written": b.numel() * 4, "repeats": 3}))
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): this build has no multi-GPU path; "
              f"run with --gpus 1.", file=sys.stderr)
        return 2
    if args.steps < 1 or args.warmup < 1:
        print("--steps and --warmup must be >= 1 (the first warm-up step is the leapfrog's half-kick prime)",
              file=sys.stderr)
        return 2

    t0 = time.time()
    p = ic.plummer(args.n, args.seed)
    log(f"generated {args.n} Plummer particles in {time.time() - t0:.1f}s")
    fp = layout.force_params(args.theta, args.eps * args.eps, 1.0, True)
    ctx = bl.Context(0)
    ctx.set_precision(bl.BH_FP64 if args.fp64 else bl.BH_FP32)

    acc_report = accuracy(ctx, p, fp, args)
    log(f"accuracy vs FP64 direct sum: {acc_report}")

    ctx.state_upload(p)
    dt = args.dt
    ctx.state_step(fp, args.max_leaf, args.group_size, args.target_size, dt / 2, dt)  # prime (main.cc:206-208)
    for _ in range(args.warmup - 1):
        ctx.state_step(fp, args.max_leaf, args.group_size, args.target_size, dt, dt)
    lat, stats = [], []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        ctx.state_step(fp, args.max_leaf, args.group_size, args.target_size, dt, dt)
        lat.append(time.perf_counter() - s0)
        stats.append(ctx.last_step_stats().as_dict())
    wall = time.perf_counter() - t_start
    ctx.state_step(fp, args.max_leaf, args.group_size, args.target_size, dt / 2, 0.0)  # closing half-kick
    log(f"timed {args.steps} steps in {wall:.3f}s")

    walk_ms = float(np.mean([s["walk_ms"] for s in stats]))
    flops = float(np.mean([FLOP_PP * s["n_pp"] + FLOP_PC * s["n_pc"] for s in stats]))
    achieved = flops / (walk_ms * 1e-3) / 1e12
    phases = {k: float(np.mean([s[k] for s in stats])) for k in ("sort_ms", "tree_ms", "targets_ms", "walk_ms")}
    out = {
        "metric": METRIC,
        "value": args.n * args.steps / wall,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.fp64 else "f32",
        "data": "synthetic",
        "config": {
            "workload": f"plummer_{args.n}",
            "n": args.n,
            "theta": args.theta,
            "eps": args.eps,
            "max_leaf": args.max_leaf,
            "group_size": args.group_size,
            "target_size": args.target_size,
            "dt": args.dt,
            "seed": args.seed,
            "walk": "fp64" if args.fp64 else "fp32 + compensated accumulation; fp64 tree, MAC, moments, kick/drift",
        },
        "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        "accuracy": acc_report,
        "phases_ms": phases,
        "interactions_per_particle": {
            "pp": float(np.mean([s["n_pp"] for s in stats])) / args.n,
            "pc": float(np.mean([s["n_pc"] for s in stats])) / args.n,
        },
        "roofline": {
            "kernel": "k_walk",
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": None,
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(p, fp, args)
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
