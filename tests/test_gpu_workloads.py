"""The batched-ensemble and serving workloads at their real size.

Ensemble: 64 independent systems, N log-uniform in [4,096, 1,048,576] (11.9M
particles), half Plummer and half Hernquist, theta cycling 0.4 / 0.5 / 0.7,
eps 0.01, the autotuned shapes of each system. One batched step equals, byte
for byte, the same step of every system alone; every system's accelerations
are checked against the FP64 direct sum on 512 of its particles, reported
per theta: the rms over all systems of a theta is within the 1e-3 bound
wherever the tuner met the bound for every size class of that theta, and
each system stays within 1.5 times the bound (or its class's error floor,
where no shape met the bound: theta 0.7 at the largest sizes), a check for
gross errors that allows for the scatter between seeds.

Serving: the first 200 requests of the 2,000-request trace (Poisson arrivals
at 200/s, N log-uniform in [1,024, 2,097,152] Plummer, 1-50 leapfrog steps,
theta 0.5, eps 0.01) through the continuous-batching server under its
200 GB cap: the footprint stays under the cap at every step and a sample of
12 completions, small to the largest, is bit-identical to the same leapfrog
run alone.
"""
import numpy as np
import pytest

from bhgpu import autotune, devmem, ensemble, ic, layout
from bhgpu import lib as bl

pytestmark = pytest.mark.gpu


def test_ensemble_64_systems_batched_equals_alone_and_accuracy(gpu):
    p, meta = ensemble.make()
    n = len(p)
    shapes = lambda m: autotune.shapes_for(m["n"], m["theta"])  # noqa: E731
    print(f"ensemble: {len(meta)} systems, {n} particles", flush=True)
    a, b, c = (devmem.DeviceBuffer(n * 64) for _ in range(3))
    a.upload(p)
    gpu.step_batch_device(a.ptr, b.ptr, ensemble.table(meta, shapes, 1.0, 0.0))
    batched = b.download(np.empty(n, layout.PARTICLE))
    offs = np.concatenate([[0], np.cumsum([m["n"] for m in meta])])
    for k, m in enumerate(meta):
        ml, gs, ts, tile = shapes(m)
        gpu.set_walk_tile(tile)
        o = int(offs[k]) * 64
        gpu.step_device(a.ptr + o, c.ptr + o, m["n"], layout.force_params(m["theta"], m["eps2"], 1.0, True), ml, gs, ts,
                        1.0, 0.0)
    gpu.set_walk_tile(128)
    alone = c.download(np.empty(n, layout.PARTICLE))
    assert batched.tobytes() == alone.tobytes()

    # a = v_out - v_in (kick 1, no drift) against the FP64 direct sum
    rng = np.random.default_rng(777)
    per_theta, worst, met_all = {}, {}, {}
    for k, m in enumerate(meta):
        lo, hi = int(offs[k]), int(offs[k + 1])
        ps, po = p[lo:hi], batched[lo:hi]
        acc = np.empty((hi - lo, 3))
        acc[po["id"]] = po["vel"] - ps["vel"][po["id"]]
        idx = np.sort(rng.choice(hi - lo, size=min(512, hi - lo), replace=False)).astype(np.uint32)
        ref = gpu.direct_sum(ps, idx, m["eps2"], 1.0)
        rel = np.linalg.norm(acc[idx] - ref, axis=1) / np.linalg.norm(ref, axis=1)
        per_theta.setdefault(m["theta"], []).append(rel)
        rms_k = float(np.sqrt(np.mean(rel ** 2)))
        worst[m["theta"]] = max(worst.get(m["theta"], 0.0), rms_k)
        cls = autotune.class_for(m["n"], m["theta"])
        met = cls is None or cls["within_bound"]
        met_all[m["theta"]] = met_all.get(m["theta"], True) and met
        limit = 1.5 * (1e-3 if met else cls["rms_floor"])
        assert rms_k <= limit, (k, m["n"], m["theta"], rms_k, limit)
    for th, rels in sorted(per_theta.items()):
        r = np.concatenate(rels)
        rms = float(np.sqrt(np.mean(r ** 2)))
        print(f"theta {th}: rms {rms:.3e}, p99 {np.percentile(r, 99):.3e}, worst system rms {worst[th]:.3e} "
              f"({len(rels)} systems x 512 particles)", flush=True)
        if met_all[th]:
            assert rms <= 1e-3, (th, rms)


def test_serving_trace_prefix(gpu):
    R = 200
    st = ic.stream_state(4242, np.arange(2000, dtype=np.uint64))
    ns = np.rint(1024 * 2048.0 ** ic.uniform(st, 1)).astype(np.int64)[:R]
    steps = (1 + np.floor(50 * ic.uniform(st, 2)).astype(np.int64))[:R]
    parts = [ic.plummer(int(m), 5000 + k) for k, m in enumerate(ns)]

    def req(k):
        ml, gs, ts, _ = autotune.shapes_for(int(ns[k]), 0.5)
        return bl.request(int(ns[k]), theta=0.5, eps2=1e-4, max_leaf=ml, group_size=gs, target_size=ts,
                          steps=int(steps[k]), dt=1e-3)

    cap = int(200e9)
    srv = bl.Server(gpu, mem_cap_bytes=cap)
    tickets, got, i, peak = {}, {}, 0, 0
    # arrivals in trace order: eight requests per server step, so requests
    # join while others run
    while len(got) < R:
        for _ in range(8):
            if i < R:
                tickets[srv.submit(parts[i], req(i))] = i
                i += 1
        s = srv.step()
        assert s.used_bytes <= cap
        peak = max(peak, s.used_bytes)
        for t, res, c in srv.poll():
            assert c.status == 0
            got[tickets[t]] = res
    srv.close()
    print(f"serving prefix: {R} requests, {int(ns.sum())} particles, peak footprint {peak / 1e9:.1f} GB", flush=True)
    order = np.argsort(ns)
    sample = sorted(set(order[np.linspace(0, R - 1, 12).astype(int)].tolist()))
    for k in sample:
        r = req(k)
        fp = layout.force_params(0.5, 1e-4, 1.0, True)
        q = parts[k]
        sched = [(1e-3 / 2, 1e-3)] + [(1e-3, 1e-3)] * (r.steps - 1) + [(1e-3 / 2, 0.0)]
        for kick, drift in sched:
            q, _ = gpu.step(q, fp, r.max_leaf, r.group_size, r.target_size, kick, drift)
        print(f"request {k}: n={r.n} steps={r.steps}", flush=True)
        assert got[k].tobytes() == q.tobytes(), k
