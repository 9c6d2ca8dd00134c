"""bh_step_device_noalloc: the step with its scratch in the library's fixed
per-GPU array gives, byte for byte, bh_step_device's result; two calls in a
row (the second smaller, so stale scratch from the first is present) both
match; sizes beyond the limit are refused before anything runs."""
import numpy as np
import pytest

import golden_io
from bhgpu import devmem, ic, layout
from bhgpu import lib as bl

pytestmark = pytest.mark.gpu


def step_both(ctx, p, fp, ml, gs, ts, kick, drift):
    n = len(p)
    src = devmem.DeviceBuffer(n * 64)
    a, b = devmem.DeviceBuffer(n * 64), devmem.DeviceBuffer(n * 64)
    src.upload(p)
    ctx.step_device(src.ptr, a.ptr, n, fp, ml, gs, ts, kick, drift)
    want = a.download(np.empty(n, layout.PARTICLE))
    ctx.step_device_noalloc(src.ptr, b.ptr, n, fp, ml, gs, ts, kick, drift)
    got = b.download(np.empty(n, layout.PARTICLE))
    return got, want


@pytest.mark.parametrize("prec", [bl.BH_FP32, bl.BH_FP64])
def test_noalloc_equals_step_device(prec):
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    with bl.Context(0) as ctx:
        ctx.set_precision(prec)
        for k, (p, ml) in enumerate(((ic.plummer(65536, 1), 16), (ic.hernquist(20000, 2), 8))):
            print(f"noalloc call {k}: n={len(p)} max_leaf={ml} precision={prec}", flush=True)
            got, want = step_both(ctx, p, fp, ml, 256, 32, 5e-4, 1e-3)
            assert golden_io.bytes_equal(got, want), f"call {k} differs from bh_step_device"
            st = ctx.last_step_stats()
            assert st.n_pp > 0 and st.n_nodes > 0


def test_noalloc_limits():
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    with bl.Context(0) as ctx:
        n = 65537
        buf = devmem.DeviceBuffer(n * 64)
        out = devmem.DeviceBuffer(n * 64)
        with pytest.raises(bl.BHError) as e:
            ctx.step_device_noalloc(buf.ptr, out.ptr, n, fp, 16, 256, 32, 1e-3, 1e-3)
        assert e.value.code == 6
        # a tree with more nodes than the fixed array holds: 30,000 coincident
        # pairs with max_leaf 1 (chains of ~20 nodes per pair)
        base = ic.random_box(30000, 3, 10.0)
        p = np.concatenate([base, base])
        p["id"] = np.arange(len(p), dtype=np.uint32)
        src = devmem.DeviceBuffer(len(p) * 64)
        src.upload(p)
        with pytest.raises(bl.BHError) as e:
            ctx.step_device_noalloc(src.ptr, out.ptr, len(p), fp, 1, 256, 32, 1e-3, 1e-3)
        assert e.value.code == 6
        # the context still steps afterwards
        q = ic.plummer(4096, 4)
        got, want = step_both(ctx, q, fp, 16, 256, 32, 1e-3, 1e-3)
        assert golden_io.bytes_equal(got, want)
