"""Resident steps build from the compact positions their previous step's
walk wrote (half the bytes of the particle records); every other path reads
the records. Bars: a resident leapfrog equals, byte for byte, the same
leapfrog stepped through bh_step_device on two buffers (which never carries),
eager and as graph replays; a value that becomes non-finite in one resident
step is reported by the next, carried or not."""
import numpy as np
import pytest

import golden_io
from bhgpu import devmem, ic, layout
from bhgpu import lib as bl

pytestmark = pytest.mark.gpu

SCHED = [(2.5e-4, 5e-4)] + [(5e-4, 5e-4)] * 4 + [(2.5e-4, 0.0)]


def device_run(ctx, p, fp):
    n = len(p)
    buf = [devmem.DeviceBuffer(n * 64) for _ in range(2)]
    buf[0].upload(p)
    cur = 0
    for kick, drift in SCHED:
        ctx.step_device(buf[cur].ptr, buf[1 - cur].ptr, n, fp, 16, 256, 64, kick, drift)
        cur = 1 - cur
    return buf[cur].download(np.empty(n, layout.PARTICLE))


@pytest.mark.parametrize("graph,prec,n", [(False, bl.BH_FP32, 30000), (True, bl.BH_FP32, 30000),
                                          (False, bl.BH_FP64, 30000), (False, bl.BH_FP32, 37)])
def test_resident_steps_equal_device_steps(graph, prec, n):
    p = ic.plummer(n, 81)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    with bl.Context(0) as ctx:
        ctx.set_graph_mode(bl.BH_GRAPH_ON if graph else bl.BH_GRAPH_OFF)
        ctx.set_precision(prec)
        print(f"carried positions graph={graph} precision={prec} n={n}", flush=True)
        want = device_run(ctx, p, fp)
        ctx.state_upload(p)
        for kick, drift in SCHED:
            ctx.state_step(fp, 16, 256, 64, kick, drift)
        assert golden_io.bytes_equal(ctx.state_download(), want)


def test_resident_steps_equal_device_steps_when_order_changes():
    """Drifts of the order of the box: every resident step reorders the
    particles completely, so the build gathers the compact copy at
    scattered positions; still byte-equal to the device steps."""
    p = ic.random_box(20000, 83, 10.0)
    p["vel"] = np.random.default_rng(84).normal(0.0, 3.0, (len(p), 3))
    fp = layout.force_params(0.6, 1e-4, 1.0, True)
    sched = [(0.0, 1.0), (1e-3, 1.0), (0.0, 1.0), (1e-3, 0.5)]
    with bl.Context(0) as ctx:
        n = len(p)
        buf = [devmem.DeviceBuffer(n * 64) for _ in range(2)]
        buf[0].upload(p)
        cur = 0
        for kick, drift in sched:
            ctx.step_device(buf[cur].ptr, buf[1 - cur].ptr, n, fp, 8, 256, 32, kick, drift)
            cur = 1 - cur
        want = buf[cur].download(np.empty(n, layout.PARTICLE))
        print("carried positions, reordering drifts", flush=True)
        ctx.state_upload(p)
        for kick, drift in sched:
            ctx.state_step(fp, 8, 256, 32, kick, drift)
        assert golden_io.bytes_equal(ctx.state_download(), want)


@pytest.mark.parametrize("expose", [False, True])
def test_carried_nonfinite_reported(expose):
    """A drift that overflows a position to infinity: the step that makes it
    succeeds (its own input is finite), the next reports BH_ERR_NONFINITE,
    from the carried flag or (state exposed in between) from the records."""
    p = ic.plummer(5000, 82)
    p["vel"][123, 0] = 1e300
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    with bl.Context(0) as ctx:
        ctx.state_upload(p)
        ctx.state_step(fp, 16, 256, 32, 0.0, 1e10)
        if expose:
            ctx.state_device_ptr()
        with pytest.raises(bl.BHError) as e:
            ctx.state_step(fp, 16, 256, 32, 0.0, 1e-3)
        assert e.value.code == 2


def test_resident_batch_equals_device_batch():
    """bh_state_step_batch carries the compact positions across a batch of
    systems too: its leapfrog equals bh_step_batch on two buffers."""
    specs = [(ic.plummer(3000, 91), 0.5, 16, 64), (ic.hernquist(12000, 92), 0.7, 8, 32),
             (ic.random_box(700, 93, 2.0, vel_spread=0.2), 0.4, 16, 16), (ic.plummer(20000, 94), 0.6, 32, 128)]
    p = np.concatenate([q for q, _, _, _ in specs])
    n = len(p)

    def table(kick, drift):
        return np.concatenate([layout.system(len(q), layout.force_params(theta, 1e-4, 1.0, True), ml, 256, ts, kick, drift)
                               for q, theta, ml, ts in specs])

    with bl.Context(0) as ctx:
        buf = [devmem.DeviceBuffer(n * 64) for _ in range(2)]
        buf[0].upload(p)
        cur = 0
        for kick, drift in SCHED:
            ctx.step_batch_device(buf[cur].ptr, buf[1 - cur].ptr, table(kick, drift))
            cur = 1 - cur
        want = buf[cur].download(np.empty(n, layout.PARTICLE))
        ctx.state_upload(p)
        for kick, drift in SCHED:
            print(f"resident batch kick={kick} drift={drift}", flush=True)
            ctx.state_step_batch(table(kick, drift))
        assert golden_io.bytes_equal(ctx.state_download(), want)
