"""Batched systems through bh_step_batch: each system of a batch gets, bit for
bit, the result of the same step on it alone (bh_step), for mixed sizes and
parameters, for batches of more than 256 systems (two system-id sort
passes), and with empty systems in the table."""
import numpy as np
import pytest

from bhgpu import ic, layout
from bhgpu import lib as bl

pytestmark = pytest.mark.gpu


def make_batch(specs):
    """specs: (particles, theta, eps2, quad, max_leaf, group_size, target_size, kick, drift)."""
    parts, table = [], []
    for p, th, e2, q, ml, gs, ts, kdt, ddt in specs:
        parts.append(p)
        table.append(layout.system(len(p), layout.force_params(th, e2, 1.0, q), ml, gs, ts, kdt, ddt))
    return np.concatenate(parts) if parts else np.zeros(0, layout.PARTICLE), np.concatenate(table)


def check_alone(gpu, specs):
    allp, table = make_batch(specs)
    print(f"batch of {len(specs)} systems, {len(allp)} particles", flush=True)
    out, t = gpu.step_batch(allp, table)
    off, bad = 0, []
    for k, (p, th, e2, q, ml, gs, ts, kdt, ddt) in enumerate(specs):
        n = len(p)
        if n:
            alone, _ = gpu.step(p, layout.force_params(th, e2, 1.0, q), ml, gs, ts, kdt, ddt)
            got = out[off:off + n]
            if got.tobytes() != alone.tobytes():
                same_order = bool(np.array_equal(got["id"], alone["id"]))
                msg = f"system {k} (n={n} theta={th} eps2={e2} ml={ml} gs={gs} ts={ts}): same order {same_order}"
                if same_order:
                    msg += (f", max |dpos| {np.abs(got['pos'] - alone['pos']).max():.3e}"
                            f", max |dvel| {np.abs(got['vel'] - alone['vel']).max():.3e}"
                            f", differing particles {int((got['vel'] != alone['vel']).any(1).sum())}")
                print(msg, flush=True)
                bad.append(k)
        off += n
    assert not bad, f"systems differing from their step alone: {bad}"
    return out, t


MIXED = [
    (1, 0.5, 1e-4, True, 8, 256, 32), (2, 0.5, 0.0, True, 8, 64, 32), (33, 0.7, 1e-4, False, 4, 16, 8),
    (700, 0.4, 1e-4, True, 16, 256, 32), (5000, 0.5, 0.0, True, 1, 64, 64), (20000, 0.7, 1e-2, True, 32, 256, 32),
    (3, 0.0, 1e-4, True, 8, 256, 32), (70000, 0.5, 1e-4, True, 16, 256, 32), (12000, 0.4, 1e-4, False, 16, 8, 1),
]


@pytest.mark.parametrize("prec", [bl.BH_FP32, bl.BH_FP64])
def test_batch_equals_alone_mixed(gpu, prec):
    specs = []
    for k, (n, th, e2, q, ml, gs, ts) in enumerate(MIXED):
        p = ic.plummer(n, 100 + k) if k % 2 == 0 else ic.hernquist(n, 100 + k)
        specs.append((p, th, e2, q, ml, gs, ts, 5e-4 * (1 + k % 3), 1e-3 * (k % 2)))
    # a deep tree (a 120-fold duplicate: a chain down to depth 21) among shallow ones
    deep = ic.random_box(400, 77, 2.0, vel_spread=0.1)
    deep["pos"][:120] = deep["pos"][5]
    specs.append((deep, 0.5, 1e-4, True, 8, 64, 32, 5e-4, 1e-3))
    gpu.set_precision(prec)
    try:
        out, t = check_alone(gpu, specs)
    finally:
        gpu.set_precision(bl.BH_FP32)
    assert t.tree_build_ms > 0 and t.traversal_ms > 0


def test_batch_many_small_systems(gpu):
    rng = np.random.default_rng(5)
    sizes = rng.integers(0, 300, size=300)
    sizes[[0, 7, 150]] = 0
    specs = [(ic.random_box(int(n), 1000 + k, 3.0, vel_spread=0.1), 0.5, 1e-4, True, 8, 64, 16, 1e-3, 1e-3)
             for k, n in enumerate(sizes)]
    check_alone(gpu, specs)


def test_batch_all_systems_one_size_identical(gpu):
    # 260 copies of one system: every copy ends identical to the system alone
    p = ic.plummer(257, 4)
    specs = [(p, 0.5, 1e-4, True, 8, 64, 32, 1e-3, 1e-3)] * 260
    allp, table = make_batch(specs)
    out, _ = gpu.step_batch(allp, table)
    alone, _ = gpu.step(p, layout.force_params(0.5, 1e-4, 1.0, True), 8, 64, 32, 1e-3, 1e-3)
    assert np.array_equal(out.reshape(260, 257).view(np.uint8).reshape(260, -1),
                          np.broadcast_to(alone.view(np.uint8), (260, alone.nbytes)))


def test_batch_errors(gpu):
    p = ic.random_box(100, 1, 1.0)
    good = layout.system(100, layout.force_params(0.5, 1e-4), 8, 64, 32, 1e-3, 1e-3)
    bad = layout.system(100, layout.force_params(-1.0, 1e-4), 8, 64, 32, 1e-3, 1e-3)
    with pytest.raises(bl.BHError):
        gpu.step_batch(np.concatenate([p, p]), np.concatenate([good, bad]))
    too_many = np.zeros(65536, layout.SYSTEM)
    too_many["group_size"] = 64
    with pytest.raises(bl.BHError):
        gpu.step_batch(np.zeros(0, layout.PARTICLE), too_many)
    empty = np.concatenate([layout.system(0, layout.force_params(), 8, 64, 32, 0, 0)] * 3)
    out, _ = gpu.step_batch(np.zeros(0, layout.PARTICLE), empty)
    assert len(out) == 0
    # the context still works after the errors
    out, _ = gpu.step_batch(p, good)
    alone, _ = gpu.step(p, layout.force_params(0.5, 1e-4), 8, 64, 32, 1e-3, 1e-3)
    assert out.tobytes() == alone.tobytes()


def test_batch_walk_tiles_equal_alone(gpu):
    """Systems of one batch with their own walk tiles each get the bytes of
    a step alone with that tile."""
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    parts = [ic.plummer(6000, 201), ic.hernquist(9000, 202), ic.plummer(4096, 203)]
    tiles = [16, 64, 0]
    table = np.concatenate([layout.system(len(q), fp, 8, 256, 32, 1e-3, 1e-3, walk_tile=t) for q, t in zip(parts, tiles)])
    out, _ = gpu.step_batch(np.concatenate(parts), table)
    off = 0
    try:
        for q, t in zip(parts, tiles):
            gpu.set_walk_tile(t or 128)
            alone, _ = gpu.step(q, fp, 8, 256, 32, 1e-3, 1e-3)
            assert out[off:off + len(q)].tobytes() == alone.tobytes(), t
            off += len(q)
    finally:
        gpu.set_walk_tile(128)
