"""Benchmark: particles/s per step of the Barnes-Hut force path on one MI355X.

Workload (the reference configuration the metric is quoted on): Plummer
sphere N = 16,777,216 (G = M = a = 1, r < 10a, seed 3), eps = 1e-3,
theta = 0.5, leaf capacity 32, leapfrog dt = 5e-4, FP32 walk with
batch sums accumulated in FP64, FP64 tree / multipoles / kick-drift. Group size
256 is the reference driver's default (main.cc:30-33); the target size is
96 instead of its 32 (measured at this workload: the fastest of the sizes
32 to 128, rms error 2.6e-4 or below for all; --target-size 32 gives the
default), and the
walk tile is the largest, 128 particles per walk unit.

A step is one step_particles call (tree build + walk + fused kick/drift) on
particles resident in HBM. The leapfrog's priming half-kick is the first
warm-up step; the K timed steps are full kick+drift steps.

Prints one JSON line. `roofline` is for the walk kernel (the dominant one):
algorithmic FP32 flops per launch (19 per particle interaction, 55 per
quadrupole-cell interaction) over its average duration from HIP events on
the library's stream. `cpu_baseline` times the FP64 oracle (a C++
restatement of the reference's grouped path) on the host's usable cores over
a bounded sample of the same workload.

Beside the headline the line carries each of the other workloads: the
65,536-particle FP64 leapfrog (`config0`), the 1M uniform cube (`config1`),
the 16M run's energy drift (`energy`), a 4,096-particle step's latency eager
and as a graph replay (`small_system_latency`), the 64-system mixed
ensemble batched and one system per call, each eager and graph-captured
(`ensemble`), and the 2,000-request serving trace (`serving`).
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

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
    """FP32 tree accelerations vs the FP64 direct sum on a seeded subset; also
    returns the accelerations (Morton order) for the tree-force comparison."""
    ctx.build_tree(p, args.max_leaf, args.group_size)
    acc, _, _, _ = ctx.accel(fp, args.target_size)
    perm = ctx.export_perm()  # Morton position -> input index
    rng = np.random.default_rng(12345)
    idx = np.sort(rng.choice(len(p), size=min(args.error_subset, len(p)), replace=False)).astype(np.uint32)
    where = np.empty(len(p), np.int64)
    where[perm] = np.arange(len(p))
    ref = ctx.direct_sum(p, idx, fp["eps2"][0], fp["G"][0])
    rms, p99, glob = rel_errors(acc[where[idx]], ref)
    return {"rms_rel": rms, "p99_rel": p99, "global_rel": glob, "subset": int(len(idx)), "reference": "FP64 direct sum"}, acc


TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "k_walk_plummer16m.md")


def profiled_traffic(args):
    """HBM bytes per walk launch at this workload from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs (calibrated on a 1 GiB copy), or
    None when the profile does not cover this configuration."""
    if args.n != 1 << 24 or args.fp64 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        if line.startswith("| `k_walk<float"):
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 4:
                return (float(cells[2]) + float(cells[3])) * 1e6
    return None


def usable_cpus():
    """CPUs this process may run on: its affinity set, capped by the cgroup
    CPU quota when one is set. Returns (usable, affinity, quota or None)."""
    try:
        aff = len(os.sched_getaffinity(0))
    except AttributeError:
        aff = os.cpu_count() or 1
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            quota = max(1, -(-int(q) // int(period)))
    except (OSError, ValueError):
        pass
    return max(1, min(aff, quota) if quota else aff), aff, quota


def host_info():
    """CPU model, and the oracle's compiler and flags (oracle/Makefile)."""
    import subprocess

    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        cxx = subprocess.run(["g++", "--version"], capture_output=True, text=True, timeout=10).stdout.splitlines()[0]
    except (OSError, IndexError, subprocess.SubprocessError):
        cxx = "g++ (version not available on this host)"
    flags = ""
    try:
        for line in open(os.path.join(REPO, "oracle", "Makefile")):
            if line.startswith("CXXFLAGS"):
                flags = line.split(":=", 1)[1].strip()
    except OSError:
        pass
    return {"cpu_model": model, "compiler": cxx, "flags": flags}


def cpu_baseline(p, fp, args, gpu_acc=None):
    """The FP64 oracle on the host's cores over a bounded sample; with the
    GPU's accelerations of the same tree (Morton order) also returns their
    error against the oracle's FP64 tree forces on the sampled groups."""
    sys.path.insert(0, REPO)
    import oracle  # the FP64 restatement, timed as the CPU baseline

    cores, affinity, quota = usable_cpus()
    o = oracle.Oracle()
    (build_ms, walk_ms, frac), oacc, mask = o.baseline_sample(p, fp, args.max_leaf, args.group_size,
                                                              args.target_size, args.cpu_stride, cores, with_acc=True)
    tree_err = None
    if gpu_acc is not None:
        rms, p99, glob = rel_errors(gpu_acc[mask], oacc[mask])
        tree_err = {"rms_rel": rms, "p99_rel": p99, "global_rel": glob, "particles": int(mask.sum()),
                    "reference": "FP64 grouped walk (oracle) on the same tree, every "
                                 f"{args.cpu_stride}th group"}
    step_s = build_ms / 1e3 + (walk_ms / 1e3) / frac
    return tree_err, {
        "value": len(p) / step_s,
        "unit": "particles/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"FP64 oracle, full tree build of all {len(p)} particles ({build_ms:.0f} ms) + grouped walk of "
                   f"every {args.cpu_stride}th group ({100 * frac:.2f}% of particles, {walk_ms:.0f} ms), walk time "
                   f"scaled to all groups"),
        "threads": {"used": cores, "affinity": affinity, "cgroup_quota": quota},
        "host": host_info(),
        "note": "baseline only: the reference's own build needs TTG/PaRSEC, which are not available",
    }


def ens_shapes(args, n, theta):
    """(max_leaf, group_size, target_size, walk_tile): the autotuned table's
    choice for (n, theta), unless fixed on the command line (--ens-max-leaf > 0)."""
    if args.ens_max_leaf > 0:
        return args.ens_max_leaf, args.ens_group_size, args.ens_target_size, 128
    from bhgpu import autotune

    return autotune.shapes_for(n, theta)


def rel_rows(acc, ref):
    return np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)


def ensemble_run(ctx, args):
    """The mixed-N ensemble (64 systems, N log-uniform in [4096, 2^20], half
    Plummer / half Hernquist, theta in {0.4, 0.5, 0.7}, eps 0.01): the
    batched path (bh_step_batch, every system in the same launches) against
    the plain path (one bh_step_device call per system), each eager and as
    graph replays, each over the leapfrog schedule of `ensemble_steps` steps
    from the same resident start."""
    from bhgpu import devmem, ensemble

    t0 = time.time()
    p, meta = ensemble.make()
    n = len(p)
    log(f"ensemble: {len(meta)} systems, {n} particles, generated in {time.time() - t0:.1f}s")
    shapes = lambda m: ens_shapes(args, m["n"], m["theta"])  # noqa: E731
    dt = 1e-3
    sched = [(dt / 2, dt)] + [(dt, dt)] * (args.ensemble_steps - 1) + [(dt / 2, 0.0)]
    tables = {kd: ensemble.table(meta, shapes, *kd) for kd in set(sched)}
    offs = np.concatenate([[0], np.cumsum([m["n"] for m in meta])])
    fps = [layout.force_params(m["theta"], m["eps2"], 1.0, True) for m in meta]
    buf = [devmem.DeviceBuffer(n * 64) for _ in range(2)]

    def run(mode, schedule):
        batched = mode.startswith("batched")
        ctx.set_graph_mode(bl.BH_GRAPH_ON if mode.endswith("graph") else bl.BH_GRAPH_OFF)
        buf[0].upload(p)
        devmem.synchronize()
        cur, lat = 0, []
        t_start = time.perf_counter()
        for kick, drift in schedule:
            s0 = time.perf_counter()
            a, b = buf[cur].ptr, buf[1 - cur].ptr
            if batched:
                ctx.step_batch_device(a, b, tables[(kick, drift)])
            else:
                for k, m in enumerate(meta):
                    o = int(offs[k]) * 64
                    ml, gs, ts, tile = shapes(m)
                    ctx.set_walk_tile(tile)
                    ctx.step_device(a + o, b + o, m["n"], fps[k], ml, gs, ts, kick, drift)
                ctx.set_walk_tile(128)
            lat.append(time.perf_counter() - s0)
            cur ^= 1
        wall = time.perf_counter() - t_start
        return wall, lat, buf[cur].download(np.empty(n, layout.PARTICLE))

    modes = ("batched", "batched_graph", "per_system", "per_system_graph")
    for mode in modes:
        run(mode, sched[:2])  # grow the context's buffers, capture both buffer directions
    res, finals = {}, {}
    for mode in modes:
        wall, lat, fin = run(mode, sched)
        finals[mode] = fin.view(np.uint8)
        res[mode] = {
            "particles_per_s": n * len(sched) / wall,
            "ms_per_step": 1e3 * wall / len(sched),
            "latency_ms": {"p50": 1e3 * float(np.percentile(lat, 50)), "p99": 1e3 * float(np.percentile(lat, 99))},
        }
        log(f"ensemble {mode}: {res[mode]}")
    graphs_held = ctx.graph_count()
    ctx.set_graph_mode(bl.BH_GRAPH_OFF)
    identical = all(np.array_equal(finals[m], finals["batched"]) for m in modes)

    # accuracy of the batched step: a = v_out - v_in from one step with
    # kick_dt = 1, drift_dt = 0, against the FP64 direct sum per system
    kd = (1.0, 0.0)
    buf[0].upload(p)
    devmem.synchronize()
    ctx.step_batch_device(buf[0].ptr, buf[1].ptr, ensemble.table(meta, shapes, *kd))
    out = buf[1].download(np.empty(n, layout.PARTICLE))
    rng = np.random.default_rng(777)
    per_theta = {}
    systems = []
    for k, m in enumerate(meta):
        lo, hi = int(offs[k]), int(offs[k + 1])
        ps, po = p[lo:hi], out[lo:hi]
        acc = np.empty((hi - lo, 3))
        acc[po["id"]] = po["vel"] - ps["vel"][po["id"]]
        idx = np.sort(rng.choice(hi - lo, size=min(args.ens_error_subset, hi - lo), replace=False)).astype(np.uint32)
        rel = rel_rows(acc[idx], ctx.direct_sum(ps, idx, m["eps2"], 1.0))
        rms = float(np.sqrt(np.mean(rel ** 2)))
        systems.append({"n": m["n"], "kind": m["kind"], "theta": m["theta"], "shapes": list(shapes(m)), "rms_rel": rms,
                        "p99_rel": float(np.percentile(rel, 99)), "within_bound": rms <= 1e-3})
        per_theta.setdefault(m["theta"], []).append(rel)
    accuracy = {"subset_per_system": args.ens_error_subset, "reference": "FP64 direct sum", "bound": 1e-3}
    for th, rels in sorted(per_theta.items()):
        r = np.concatenate(rels)
        worst = max(s["rms_rel"] for s in systems if s["theta"] == th)
        accuracy[f"theta_{th:g}"] = {"rms_rel": float(np.sqrt(np.mean(r ** 2))), "p99_rel": float(np.percentile(r, 99)),
                                     "worst_system_rms_rel": worst,
                                     "systems_within_bound": sum(s["within_bound"] for s in systems if s["theta"] == th),
                                     "systems": sum(1 for s in systems if s["theta"] == th)}
    allr = np.concatenate([np.concatenate(v) for v in per_theta.values()])
    accuracy["all"] = {"rms_rel": float(np.sqrt(np.mean(allr ** 2))), "p99_rel": float(np.percentile(allr, 99)),
                       "worst_system_rms_rel": max(s["rms_rel"] for s in systems)}
    pps = lambda m: res[m]["particles_per_s"]  # noqa: E731
    return {
        "workload": f"ensemble_{len(meta)}x_loguniform_4096_1048576",
        "systems": len(meta),
        "particles": int(n),
        "evaluations": len(sched),
        "shapes": ("autotuned per system (bhgpu/tuning.json): " +
                   ", ".join(sorted({str(shapes(m)) for m in meta}))) if args.ens_max_leaf == 0 else
                  {"max_leaf": args.ens_max_leaf, "group_size": args.ens_group_size,
                   "target_size": args.ens_target_size},
        **res,
        "speedup": pps("batched") / pps("per_system"),
        "speedup_vs_graphed_per_system": max(pps("batched"), pps("batched_graph")) / pps("per_system_graph"),
        "graphs_held": graphs_held,
        "bit_identical_across_modes": identical,
        "accuracy": accuracy,
        "per_system_accuracy": systems,
    }


def small_latency_run(ctx, n=4096, reps=50):
    """One 4,096-particle Plummer step (leaf 16, group 256, target 32,
    theta 0.5, eps 0.01) on device buffers: host wall time per call and the
    device time of the step, eager and as a graph replay."""
    from bhgpu import devmem

    p = ic.plummer(n, 81)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    a, b = devmem.DeviceBuffer(n * 64), devmem.DeviceBuffer(n * 64)
    a.upload(p)
    out = {"workload": f"plummer_{n}", "n": n, "calls": reps}
    for name, mode in (("eager", bl.BH_GRAPH_OFF), ("graph", bl.BH_GRAPH_ON)):
        ctx.set_graph_mode(mode)
        for _ in range(3):
            ctx.step_device(a.ptr, b.ptr, n, fp, 16, 256, 32, 1e-3, 1e-3)
        lat, dev = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            ctx.step_device(a.ptr, b.ptr, n, fp, 16, 256, 32, 1e-3, 1e-3)
            lat.append(1e3 * (time.perf_counter() - t0))
            dev.append(ctx.last_step_stats().step_ms)
        out[name] = {"host_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
                     "device_ms_p50": float(np.percentile(dev, 50)),
                     "particles_per_s": n / (float(np.percentile(lat, 50)) * 1e-3)}
    ctx.set_graph_mode(bl.BH_GRAPH_OFF)
    return out


def config0_run(ctx, args):
    """Plummer N = 65,536 (seed 1), FP64 walk, eps 0.01, theta 0.5, leaf 16,
    10 leapfrog steps of dt 1e-3 (11 force evaluations) on resident
    particles; accelerations of the initial state against the FP64 direct
    sum over all particles."""
    n = 65536
    p = ic.plummer(n, 1)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    ml, gs, ts, dt, steps = 16, 256, 32, 1e-3, 10
    ctx.set_precision(bl.BH_FP64)
    try:
        ctx.build_tree(p, ml, gs)
        acc, _, _, _ = ctx.accel(fp, ts)
        perm = ctx.export_perm()
        rel = rel_rows(acc, ctx.direct_sum(p[perm], None, 1e-4, 1.0))
        sched = [(dt / 2, dt)] + [(dt, dt)] * (steps - 1) + [(dt / 2, 0.0)]
        ctx.state_upload(p)
        ctx.state_step(fp, ml, gs, ts, 0.0, 0.0)  # warm-up (grows buffers), not part of the run
        ctx.state_upload(p)
        t0 = time.perf_counter()
        for kick, drift in sched:
            ctx.state_step(fp, ml, gs, ts, kick, drift)
        wall = time.perf_counter() - t0
    finally:
        ctx.set_precision(bl.BH_FP64 if args.fp64 else bl.BH_FP32)
    return {"workload": "plummer_65536_fp64_leapfrog10", "n": n, "evaluations": len(sched), "walk": "fp64",
            "particles_per_s": n * len(sched) / wall, "ms_per_evaluation": 1e3 * wall / len(sched),
            "accuracy": {"rms_rel": float(np.sqrt(np.mean(rel ** 2))), "p99_rel": float(np.percentile(rel, 99)),
                         "subset": n, "reference": "FP64 direct sum over all particles (initial state)"}}


def config1_run(ctx, args):
    """Uniform cube N = 2^20 in [0,1)^3 (seed 2), eps 1e-3, theta 0.6, leaf 8
    (the reference driver's default), one tree build + force evaluation;
    FP32 and FP64 walk errors against the FP64 direct sum on a random
    4,096-particle subset."""
    n = 1 << 20
    p = ic.uniform_cube(n, 2)
    fp = layout.force_params(0.6, 1e-6, 1.0, True)
    ml, gs, ts = 8, 256, 32
    rng = np.random.default_rng(2)
    idx = np.sort(rng.choice(n, size=4096, replace=False)).astype(np.uint32)
    ref = ctx.direct_sum(p, idx, 1e-6, 1.0)
    out = {"workload": "uniform_1048576_single_eval", "n": n, "theta": 0.6, "eps": 1e-3, "max_leaf": ml}
    for name, prec in (("fp32", bl.BH_FP32), ("fp64", bl.BH_FP64)):
        ctx.set_precision(prec)
        ctx.build_tree(p, ml, gs)
        acc, _, _, _ = ctx.accel(fp, ts)
        where = np.empty(n, np.int64)
        where[ctx.export_perm()] = np.arange(n)
        rel = rel_rows(acc[where[idx]], ref)
        ctx.state_upload(p)
        ctx.state_step(fp, ml, gs, ts, 0.0, 0.0)  # warm-up
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            ctx.state_step(fp, ml, gs, ts, 0.0, 0.0)  # build + evaluation, no kick or drift
            times.append(time.perf_counter() - t0)
        t = float(np.median(times))
        out[name] = {"particles_per_s": n / t, "ms_per_evaluation": 1e3 * t,
                     "rms_rel": float(np.sqrt(np.mean(rel ** 2))), "p99_rel": float(np.percentile(rel, 99))}
    ctx.set_precision(bl.BH_FP64 if args.fp64 else bl.BH_FP32)
    out["accuracy_reference"] = "FP64 direct sum on 4,096 random particles"
    return out


def serving_run(ctx, args):
    """A prefix of the serving trace: requests with Poisson arrivals at
    `serve_rate`/s, N log-uniform in [1024, 2^21] Plummer systems (request k
    from seed 5000 + k), 1-50 leapfrog steps, theta 0.5, eps 0.01, replayed
    against the server clock through the continuous-batching server
    (admission by footprint under `serve_cap_gb`). Latency runs from a
    request's scheduled arrival to its completion."""
    R = args.serve_requests
    st = ic.stream_state(4242, np.arange(R, dtype=np.uint64))
    arrivals = np.cumsum(-np.log1p(-ic.uniform(st, 0)) / args.serve_rate)
    ns = np.rint(1024 * 2048.0 ** ic.uniform(st, 1)).astype(np.int64)
    steps = 1 + np.floor(50 * ic.uniform(st, 2)).astype(np.int64)
    t0 = time.time()
    # each request from its own seed, generated on the host's cores
    with ThreadPoolExecutor(max_workers=usable_cpus()[0]) as ex:
        parts = list(ex.map(lambda kn: ic.plummer(int(kn[1]), 5000 + kn[0]), enumerate(ns)))
    log(f"serving: {R} requests, {int(ns.sum())} particles, generated in {time.time() - t0:.1f}s")
    def req(n, k):
        ml, gs, ts, _ = ens_shapes(args, n, 0.5)  # requests use the context's walk tile
        return bl.request(n, theta=0.5, eps2=1e-4, max_leaf=ml, group_size=gs, target_size=ts, steps=k, dt=1e-3)

    reqs = [req(int(n), int(k)) for n, k in zip(ns, steps)]
    srv = bl.Server(ctx, mem_cap_bytes=int(args.serve_cap_gb * 1e9))
    # warm the context's buffers with the largest request alone
    big = int(np.argmax(ns))
    srv.submit(parts[big], req(int(ns[big]), 1))
    while srv.pending() != (0, 0):
        srv.step()
    srv.poll()
    ticket_of = {}
    lat, batch_sizes, batch_parts, step_ms, used = [], [], [], [], []
    i, done = 0, 0
    start = srv.clock()
    while done < R:
        now = srv.clock() - start
        while i < R and arrivals[i] <= now:
            ticket_of[srv.submit(parts[i], reqs[i])] = i
            i += 1
        if srv.pending() == (0, 0):
            if i >= R:
                break
            time.sleep(max(0.0, min(arrivals[i] - (srv.clock() - start), 0.05)))
            continue
        stt = srv.step()
        if stt.stepped_particles:
            batch_sizes.append(stt.stepped_systems)
            batch_parts.append(stt.stepped_particles)
            step_ms.append(stt.step_ms)
            used.append(stt.used_bytes)
        for t, _, c in srv.poll():
            k = ticket_of.pop(t)
            lat.append(c.done_s - (start + arrivals[k]))
            done += 1
    makespan = srv.clock() - start
    srv.close()
    work = float((ns * (steps + 1)).sum())
    return {
        "workload": f"serving_trace_{R}req_poisson{args.serve_rate:g}_loguniform_1024_2097152",
        "requests": R,
        "particle_evaluations": work,
        "makespan_s": makespan,
        "particles_per_s": work / makespan,
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "max": float(np.max(lat))},
        "batch": {"steps": len(step_ms), "systems_mean": float(np.mean(batch_sizes)),
                  "particles_mean": float(np.mean(batch_parts)), "step_ms_p50": float(np.percentile(step_ms, 50)),
                  "step_ms_p99": float(np.percentile(step_ms, 99)), "peak_footprint_gb": max(used) / 1e9},
        "admission_cap_gb": args.serve_cap_gb,
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
    ap.add_argument("--target-size", type=int, default=96)
    ap.add_argument("--walk-tile", type=int, default=128, help="particles per walk unit (16-128)")
    ap.add_argument("--fp64", action="store_true", help="FP64 walk arithmetic")
    ap.add_argument("--error-subset", type=int, default=4096)
    ap.add_argument("--cpu-stride", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the mixed-N ensemble comparison")
    ap.add_argument("--ensemble-only", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=10)
    ap.add_argument("--ens-max-leaf", type=int, default=0, help="0: autotuned shapes per system")
    ap.add_argument("--ens-group-size", type=int, default=256)
    ap.add_argument("--ens-target-size", type=int, default=32)
    ap.add_argument("--ens-error-subset", type=int, default=512)
    ap.add_argument("--serve-requests", type=int, default=2000, help="serving-trace requests (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-cap-gb", type=float, default=200.0)
    ap.add_argument("--serving-only", action="store_true")
    ap.add_argument("--no-energy", action="store_true", help="skip the 16M run's O(N^2) energy drift")
    ap.add_argument("--no-configs", action="store_true", help="skip the 65k, 1M and 4,096-particle workloads")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device three times, to calibrate HBM byte counters")
    args = ap.parse_args()

    if args.copy_calibration:
        from bhgpu import devmem

        nbytes = args.copy_calibration << 20
        a, b = devmem.DeviceBuffer(nbytes), devmem.DeviceBuffer(nbytes)
        a.upload(np.ones(nbytes // 4, np.float32))
        for _ in range(3):
            b.copy_from(a)
        devmem.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": nbytes, "bytes_written": nbytes, "repeats": 3}))
        return 0

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} requested): this build has no multi-GPU path; "
              f"run with --gpus 1.", file=sys.stderr)
        return 2
    if args.steps < 1 or args.warmup < 1:
        print("--steps and --warmup must be >= 1 (the first warm-up step is the leapfrog's half-kick prime)",
              file=sys.stderr)
        return 2

    if args.serving_only:
        ctx = bl.Context(0)
        print(json.dumps(serving_run(ctx, args)), flush=True)
        ctx.close()
        return 0
    if args.ensemble_only:
        ctx = bl.Context(0)
        print(json.dumps(ensemble_run(ctx, args)), flush=True)
        ctx.close()
        return 0

    t0 = time.time()
    p = ic.plummer(args.n, args.seed)
    log(f"generated {args.n} Plummer particles in {time.time() - t0:.1f}s")
    fp = layout.force_params(args.theta, args.eps * args.eps, 1.0, True)
    ctx = bl.Context(0)
    ctx.set_precision(bl.BH_FP64 if args.fp64 else bl.BH_FP32)
    ctx.set_walk_tile(args.walk_tile)

    acc_report, gpu_acc = accuracy(ctx, p, fp, args)
    log(f"accuracy vs FP64 direct sum: {acc_report}")
    energy = None
    if not args.no_energy:
        t0 = time.time()
        ek0, ep0 = ctx.energies(p)
        energy = {"kinetic_start": ek0, "potential_start": ep0, "potential_s": time.time() - t0}
        log(f"energy at start: {energy}")

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
    if energy is not None:
        # main.cc:115-121: total energy after the run against the start; the
        # closing half-kick brings velocities back in step with positions
        ek1, ep1 = ctx.energies(ctx.state_download())
        e0, e1 = energy["kinetic_start"] + energy["potential_start"], ek1 + ep1
        energy.update({"kinetic_end": ek1, "potential_end": ep1, "total_start": e0, "total_end": e1,
                       "drift_rel": (e1 - e0) / abs(e0), "leapfrog_steps": args.warmup + args.steps,
                       "dt": dt, "potential": "unsoftened FP64 pair sum over all N(N-1)/2 pairs (integrate.h:35-43)"})
        log(f"energy: {energy}")

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
            "walk_tile": args.walk_tile,
            "dt": args.dt,
            "seed": args.seed,
            "walk": "fp64" if args.fp64 else "fp32 interactions, batch sums accumulated in fp64; fp64 tree, MAC, moments, kick/drift",
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
            "traffic": profiled_traffic(args),
            "traffic_source": "profiles/k_walk_plummer16m.md (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0, "
                              "factors from a 1 GiB device copy)",
        },
    }
    if energy is not None:
        out["energy"] = energy
    if not args.no_configs:
        out["config0"] = config0_run(ctx, args)
        log(f"config0: {out['config0']}")
        out["config1"] = config1_run(ctx, args)
        log(f"config1: {out['config1']}")
        out["small_system_latency"] = small_latency_run(ctx)
        log(f"small system: {out['small_system_latency']}")
    if not args.no_ensemble:
        del p
        out["ensemble"] = ensemble_run(ctx, args)
        p = None
    if args.serve_requests > 0:
        out["serving"] = serving_run(ctx, args)
        log(f"serving: {out['serving']}")
    if not args.no_cpu_baseline:
        if p is None:
            p = ic.plummer(args.n, args.seed)
        tree_err, out["cpu_baseline"] = cpu_baseline(p, fp, args, gpu_acc)
        out["accuracy_vs_tree_fp64"] = tree_err
        log(f"accuracy vs the FP64 tree forces: {tree_err}")
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
