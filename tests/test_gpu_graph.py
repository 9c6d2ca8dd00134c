"""Host-sync-free steps and their HIP-graph replays.

A step waits on the host once, at its end; everything in between (node
counts, tree levels, walk units) is read on the device. That makes the
launch sequence capturable: with graph mode on, a step is captured once and
replayed. Bars: every replay is byte-identical to the eager step on the same
input; replays are interleaved with eager work and host synchronisation (a
graph that re-read stale state would show here); a tree that outgrows the
planned node arrays or levels is run again and still matches the oracle bit
for bit.
"""
import numpy as np
import pytest

import golden_io
from bhgpu import devmem, ic, layout
from bhgpu import lib as bl

pytestmark = pytest.mark.gpu


def leapfrog_device(ctx, p, fp, ml, gs, ts, dt, steps, graph):
    """main.cc:199-246's schedule on two device buffers; returns every
    evaluation's output and the stats of each."""
    n = len(p)
    buf = [devmem.DeviceBuffer(n * 64) for _ in range(2)]
    buf[0].upload(p)
    sched = [(dt / 2, dt)] + [(dt, dt)] * (steps - 1) + [(dt / 2, 0.0)]
    ctx.set_graph_mode(bl.BH_GRAPH_ON if graph else bl.BH_GRAPH_OFF)
    outs, stats = [], []
    try:
        cur = 0
        for k, (kick, drift) in enumerate(sched):
            ctx.step_device(buf[cur].ptr, buf[1 - cur].ptr, n, fp, ml, gs, ts, kick, drift)
            stats.append(ctx.last_step_stats().as_dict())
            outs.append(buf[1 - cur].download(np.empty(n, layout.PARTICLE)))
            # eager work of the program's own between replays
            ctx.direct_sum(p[:64], None, 1e-4)
            devmem.synchronize()
            cur ^= 1
    finally:
        ctx.set_graph_mode(bl.BH_GRAPH_OFF)
    return outs, stats


@pytest.mark.parametrize("kind,n,ml", [("plummer", 4096, 16), ("hernquist", 30000, 8), ("uniform_cube", 70000, 32)])
def test_graph_replays_match_eager(gpu, kind, n, ml):
    p = getattr(ic, kind)(n, 31)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    print(f"graph vs eager: {kind} n={n} max_leaf={ml}", flush=True)
    eager, est = leapfrog_device(gpu, p, fp, ml, 256, 32, 1e-3, 5, graph=False)
    graph, gst = leapfrog_device(gpu, p, fp, ml, 256, 32, 1e-3, 5, graph=True)
    assert all(s["graph"] == 0 for s in est)
    assert all(s["graph"] == 1 for s in gst)
    for k, (a, b) in enumerate(zip(eager, graph)):
        assert golden_io.bytes_equal(a, b), f"evaluation {k} differs between the graph replay and the eager step"
    for a, b in zip(est, gst):
        assert (a["n_pp"], a["n_pc"], a["n_nodes"], a["n_units"]) == (b["n_pp"], b["n_pc"], b["n_nodes"], b["n_units"])


def test_graph_batch_replays_with_changing_table(gpu):
    """One batch replayed with a different kick/drift and theta per step:
    the table is re-read on every replay."""
    specs = [(ic.plummer(3000, 41), 0.5, 16), (ic.hernquist(9000, 42), 0.7, 8), (ic.plummer(513, 43), 0.4, 4)]
    p = np.concatenate([s[0] for s in specs])
    n = len(p)

    def table(step):
        return np.concatenate([layout.system(len(q), layout.force_params(th + 0.05 * (step % 2), 1e-4, 1.0, True), ml, 256,
                                             32, 1e-3 * (1 + step), 2e-3 * (step % 3)) for q, th, ml in specs])

    def run(graph):
        buf = [devmem.DeviceBuffer(n * 64) for _ in range(2)]
        buf[0].upload(p)
        gpu.set_graph_mode(bl.BH_GRAPH_ON if graph else bl.BH_GRAPH_OFF)
        outs = []
        try:
            cur = 0
            for step in range(4):
                gpu.step_batch_device(buf[cur].ptr, buf[1 - cur].ptr, table(step))
                outs.append(buf[1 - cur].download(np.empty(n, layout.PARTICLE)))
                devmem.synchronize()
                cur ^= 1
            held = gpu.graph_count()
        finally:
            gpu.set_graph_mode(bl.BH_GRAPH_OFF)
        return outs, held

    eager, _ = run(False)
    graph, held = run(True)
    assert held >= 1
    for k, (a, b) in enumerate(zip(eager, graph)):
        assert golden_io.bytes_equal(a, b), f"step {k}"


def test_graph_cache_many_systems(gpu):
    """Systems stepped one after another through the graph cache (the
    per-system path of the ensemble): each replay equals its eager step."""
    systems = [ic.plummer(n, 50 + k) for k, n in enumerate((4096, 20000, 4096 + 17, 9000))]
    fp = layout.force_params(0.5, 1e-4, 1.0, True)

    def run(graph):
        bufs = [(devmem.DeviceBuffer(len(q) * 64), devmem.DeviceBuffer(len(q) * 64)) for q in systems]
        for (a, _), q in zip(bufs, systems):
            a.upload(q)
        gpu.set_graph_mode(bl.BH_GRAPH_ON if graph else bl.BH_GRAPH_OFF)
        outs = []
        try:
            for rnd in range(3):
                for (a, b), q in zip(bufs, systems):
                    src, dst = (a, b) if rnd % 2 == 0 else (b, a)
                    gpu.step_device(src.ptr, dst.ptr, len(q), fp, 16, 256, 32, 1e-3, 1e-3)
                    outs.append(dst.download(np.empty(len(q), layout.PARTICLE)))
        finally:
            gpu.set_graph_mode(bl.BH_GRAPH_OFF)
        return outs

    for k, (a, b) in enumerate(zip(run(False), run(True))):
        assert golden_io.bytes_equal(a, b), f"call {k}"


def test_tree_outgrows_planned_nodes(orc):
    """Pairs of coincident particles with max_leaf 1 make chains down to
    depth 21: ~10 nodes per particle, far above the planned capacity. The
    step runs again with room for them; tree and step match the oracle."""
    with bl.Context(0) as ctx:
        base = ic.random_box(1500, 61, 10.0)
        p = np.concatenate([base, base])
        p["id"] = np.arange(len(p), dtype=np.uint32)
        print(f"coincident pairs n={len(p)} max_leaf=1", flush=True)
        ctx.build_tree(p, 1, 64)
        ge = ctx.export()
        oe = orc.build(p, 1, 64).export()
        for k in ("keys", "perm", "nodes", "center", "iab", "groups"):
            assert golden_io.bytes_equal(ge[k], oe[k]), k
        assert len(ge["nodes"]) > 8 * len(p)
        fresh = bl.Context(0)
        try:
            fresh.set_precision(bl.BH_FP64)
            fp = layout.force_params(0.5, 1e-4, 1.0, True)
            out, _ = fresh.step(p, fp, 1, 64, 32, 1e-3, 1e-3)
            assert fresh.last_step_stats().reruns >= 1
            ref, _ = orc.step(p, fp, 1, 64, 32, 1e-3, 1e-3)
            assert np.array_equal(out["id"], ref["id"])
            assert np.abs(out["vel"] - ref["vel"]).max() <= 1e-10 * np.abs(ref["vel"]).max()
        finally:
            fresh.close()


def test_tree_deeper_than_planned(orc):
    """A context that has seen only shallow trees plans few upward levels
    and the shallow walk stack; a deep tree then runs again with all of
    them, and its step equals the one on a fresh context."""
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    deep = ic.random_box(3000, 62, 10.0)
    deep["pos"][:200] = deep["pos"][7]  # a 200-fold duplicate: a chain to depth 21
    with bl.Context(0) as ctx, bl.Context(0) as fresh:
        ctx.step(ic.uniform_cube(4096, 63), fp, 16, 256, 32, 1e-3, 1e-3)
        assert ctx.last_step_stats().reruns == 0
        got, _ = ctx.step(deep, fp, 8, 256, 32, 1e-3, 1e-3)
        assert ctx.last_step_stats().reruns >= 1
        want, _ = fresh.step(deep, fp, 8, 256, 32, 1e-3, 1e-3)
        assert golden_io.bytes_equal(got, want)
        ref, _ = orc.step(deep, fp, 8, 256, 32, 1e-3, 1e-3)
        assert np.array_equal(got["id"], ref["id"])


def test_resident_state_survives_host_calls(gpu):
    """bh_step and bh_build_tree use their own buffers: a resident leapfrog
    interleaved with them advances exactly as one run alone; stepping a
    context with no resident state is an error."""
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    p = ic.plummer(5000, 71)
    other = ic.hernquist(7000, 72)
    with bl.Context(0) as lone, bl.Context(0) as mixed:
        with pytest.raises(bl.BHError) as e:
            mixed.state_step(fp, 16, 256, 32, 1e-3, 1e-3)
        assert e.value.code == 1
        lone.state_upload(p)
        mixed.state_upload(p)
        for k in range(3):
            lone.state_step(fp, 16, 256, 32, 1e-3, 1e-3)
            mixed.state_step(fp, 16, 256, 32, 1e-3, 1e-3)
            mixed.step(other, fp, 8, 256, 32, 1e-3, 1e-3)
            mixed.build_tree(other, 8, 64)
        assert golden_io.bytes_equal(lone.state_download(), mixed.state_download())


def test_one_sync_latency_small_system(gpu):
    """A 4,096-particle step: graph replays are not slower than eager steps
    and both report their device time (printed for the record)."""
    import time

    p = ic.plummer(4096, 81)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    n = len(p)
    a, b = devmem.DeviceBuffer(n * 64), devmem.DeviceBuffer(n * 64)
    a.upload(p)
    res = {}
    for mode in (bl.BH_GRAPH_OFF, bl.BH_GRAPH_ON):
        gpu.set_graph_mode(mode)
        for _ in range(3):
            gpu.step_device(a.ptr, b.ptr, n, fp, 16, 256, 32, 1e-3, 1e-3)
        lat = []
        for _ in range(20):
            t0 = time.perf_counter()
            gpu.step_device(a.ptr, b.ptr, n, fp, 16, 256, 32, 1e-3, 1e-3)
            lat.append(time.perf_counter() - t0)
        res[mode] = float(np.median(lat)) * 1e3
    gpu.set_graph_mode(bl.BH_GRAPH_OFF)
    print(f"4096-particle step, host wall: eager {res[bl.BH_GRAPH_OFF]:.3f} ms, graph {res[bl.BH_GRAPH_ON]:.3f} ms",
          flush=True)
    assert res[bl.BH_GRAPH_ON] < 1.5 * res[bl.BH_GRAPH_OFF]
