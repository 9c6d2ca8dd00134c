"""The continuous-batching server: requests admitted while others run, under
a memory cap, each finishing bit-identical to its leapfrog run alone."""
import numpy as np
import pytest

from bhgpu import ic, layout
from bhgpu import lib as bl

pytestmark = pytest.mark.gpu


def alone(gpu, p, req):
    fp = layout.force_params(req.fp.theta, req.fp.eps2, req.fp.G, bool(req.fp.use_quadrupole))
    dt = req.dt
    sched = [(dt / 2, dt)] + [(dt, dt)] * (req.steps - 1) + [(dt / 2, 0.0)]
    for kick, drift in sched:
        p, _ = gpu.step(p, fp, req.max_leaf, req.group_size, req.target_size, kick, drift)
    return p


def drain(srv, max_steps=1000):
    got = {}
    for _ in range(max_steps):
        for t, res, c in srv.poll():
            assert c.status == 0 and c.done_s >= c.admit_s >= c.submit_s
            got[t] = (res, c)
        pend, act = srv.pending()
        if pend == 0 and act == 0:
            break
        srv.step()
    for t, res, c in srv.poll():
        got[t] = (res, c)
    return got


def test_server_results_equal_alone(gpu):
    srv = bl.Server(gpu)
    specs = [
        (ic.plummer(3000, 1), bl.request(3000, 0.5, 1e-4, steps=3, max_leaf=16)),
        (ic.hernquist(20000, 2), bl.request(20000, 0.7, 1e-4, steps=1, max_leaf=8, group_size=64)),
        (ic.random_box(1, 3, 1.0), bl.request(1, 0.5, 0.0, steps=2)),
        (ic.plummer(70000, 4), bl.request(70000, 0.4, 1e-4, steps=4, max_leaf=32, target_size=64)),
        (ic.random_box(500, 5, 2.0, vel_spread=0.3), bl.request(500, 0.5, 0.0, use_quadrupole=False, steps=5)),
    ]
    tickets = {}
    # two now, one after a step, the rest while those run
    for k in (0, 1):
        tickets[srv.submit(*specs[k])] = k
    srv.step()
    tickets[srv.submit(*specs[2])] = 2
    srv.step()
    srv.step()
    for k in (3, 4):
        tickets[srv.submit(*specs[k])] = k
    got = drain(srv)
    assert sorted(got) == sorted(tickets)
    for t, k in tickets.items():
        res, c = got[t]
        p, req = specs[k]
        print(f"request {k} n={req.n} steps={req.steps}", flush=True)
        assert c.evaluations == req.steps + 1
        assert res.tobytes() == alone(gpu, p, req).tobytes(), k
    srv.close()


def test_server_memory_cap_serialises(gpu):
    reqs = [bl.request(4000 + 1000 * k, 0.5, 1e-4, steps=2) for k in range(4)]
    cap = max(bl.Server(gpu).footprint(r) for r in reqs) + 1
    srv = bl.Server(gpu, mem_cap_bytes=cap)
    ps = [ic.plummer(r.n, 10 + k) for k, r in enumerate(reqs)]
    tickets = [srv.submit(p, r) for p, r in zip(ps, reqs)]
    active_max = 0
    done = {}
    for _ in range(100):
        st = srv.step()
        active_max = max(active_max, st.active_systems)
        assert st.used_bytes <= cap
        for t, res, c in srv.poll():
            done[t] = (res, c)
        if len(done) == len(reqs):
            break
    assert len(done) == len(reqs) and active_max == 1
    # first come, first served
    order = sorted(done, key=lambda t: done[t][1].done_s)
    assert order == tickets
    for t, p, r in zip(tickets, ps, reqs):
        assert done[t][0].tobytes() == alone(gpu, p, r).tobytes()
    srv.close()


def test_server_rejects_bad_requests(gpu):
    srv = bl.Server(gpu, mem_cap_bytes=10 ** 6)
    p = ic.plummer(100000, 1)
    with pytest.raises(bl.BHError):
        srv.submit(p, bl.request(len(p), steps=1))  # footprint above the cap
    with pytest.raises(bl.BHError):
        srv.submit(p[:10], bl.request(10, steps=0))
    with pytest.raises(bl.BHError):
        srv.submit(p[:10], bl.request(10, theta=-1.0, steps=1))
    st = srv.step()  # nothing queued: a no-op
    assert st.active_systems == 0 and st.completed == 0
    srv.close()


def test_server_refuses_nonfinite_request(gpu):
    """A request holding a NaN is refused at submission (BH_ERR_NONFINITE);
    the requests around it complete bit-identical to their runs alone."""
    srv = bl.Server(gpu)
    good = [(ic.plummer(2000, 21), bl.request(2000, 0.5, 1e-4, steps=2, max_leaf=16)),
            (ic.hernquist(5000, 22), bl.request(5000, 0.6, 1e-4, steps=3, max_leaf=8))]
    bad = ic.plummer(3000, 23)
    bad["vel"][17, 1] = np.nan
    tickets = {srv.submit(*good[0]): 0}
    with pytest.raises(bl.BHError) as e:
        srv.submit(bad, bl.request(3000, 0.5, 1e-4, steps=2))
    assert e.value.code == 2
    srv.step()
    tickets[srv.submit(*good[1])] = 1
    got = drain(srv)
    assert sorted(got) == sorted(tickets)
    for t, k in tickets.items():
        p, req = good[k]
        assert got[t][0].tobytes() == alone(gpu, p, req).tobytes(), k
    srv.close()
