"""Full-size checks on the GPU: the uniform 1M configuration against the
oracle (keys and permutation bit-exact, accuracy vs the FP64 direct sum) and
the 16M Plummer configuration through size-independent properties."""
import numpy as np
import pytest

from bhgpu import ic, layout
from bhgpu import lib as bl

pytestmark = pytest.mark.gpu


def rel_errors(test, ref):
    num = np.linalg.norm(test - ref, axis=1)
    den = np.linalg.norm(ref, axis=1)
    rel = num / den
    return float(np.sqrt(np.mean(rel ** 2))), float(np.percentile(rel, 99))


def subset_errors(gpu, p, fp, ts, m=4096, seed=0):
    acc = gpu.accel(fp, ts)[0]
    perm = gpu.export_perm()
    where = np.empty(len(p), np.int64)
    where[perm] = np.arange(len(p))
    idx = np.sort(np.random.default_rng(seed).choice(len(p), m, replace=False)).astype(np.uint32)
    ref = gpu.direct_sum(p, idx, fp["eps2"][0], fp["G"][0])
    return rel_errors(acc[where[idx]], ref)


def test_uniform_1m(gpu, orc):
    # uniform cube N = 1,048,576, eps = 1e-3, theta = 0.6 (leaf 8: the driver default)
    n = 1 << 20
    p = ic.uniform_cube(n, 2)
    fp = layout.force_params(0.6, 1e-6, 1.0, True)
    print("uniform 1M: build", flush=True)
    gpu.build_tree(p, 8, 256)
    e = gpu.export()
    o = orc.build(p, 8, 256).export()
    for k in ("keys", "perm", "nodes", "groups"):
        assert np.array_equal(e[k].view(np.uint8), o[k].view(np.uint8)), k
    print("uniform 1M: accuracy", flush=True)
    rms, p99 = subset_errors(gpu, p, fp, 32)
    assert rms < 1e-3 and p99 < 5e-3


def test_plummer_16m_properties(gpu):
    n = 1 << 24
    p = ic.plummer(n, 3)
    fp = layout.force_params(0.5, 1e-6, 1.0, True)
    print("plummer 16M: build", flush=True)
    gpu.build_tree(p, 32, 256)
    perm = gpu.export_perm()
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.uint32))  # a permutation
    m, g, dom = gpu.tree_sizes()
    assert m > n // 32 and g > 0
    print("plummer 16M: accuracy", flush=True)
    rms, p99 = subset_errors(gpu, p, fp, 32)
    assert rms <= 1e-3, rms
    # one resident step conserves ids and mass and returns Morton order
    gpu.state_upload(p)
    gpu.state_step(fp, 32, 256, 32, 2.5e-4, 5e-4)
    q = gpu.state_download()
    assert np.array_equal(np.sort(q["id"]), np.arange(n, dtype=np.uint32))
    assert q["mass"].sum() == pytest.approx(1.0, rel=1e-12)
    st = gpu.last_step_stats()
    assert st.n_pp > 0 and st.n_pc > 0 and st.walk_ms > 0


def test_plummer_65k_fp64_leapfrog(gpu, orc):
    """The CPU-runnable configuration in full: Plummer N = 65,536 (seed 1),
    eps = 0.01, theta = 0.5, leaf 16, ten leapfrog steps of dt = 1e-3
    (main.cc:199-246, eleven force evaluations), FP64 walk, against the
    oracle step by step; then the final forces against the FP64 direct sum."""
    p = ic.plummer(1 << 16, 1)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    dt, steps = 1e-3, 10
    gpu.set_precision(bl.BH_FP64)
    try:
        gpu.state_upload(p)
        q = p
        for k, d in [(dt / 2, dt)] + [(dt, dt)] * (steps - 1) + [(dt / 2, 0.0)]:
            print(f"plummer 65k fp64: step kick={k} drift={d}", flush=True)
            gpu.state_step(fp, 16, 256, 32, k, d)
            q, _ = orc.step(q, fp, 16, 256, 32, k, d)
        out = gpu.state_download()
        assert np.array_equal(out["id"], q["id"])
        assert np.abs(out["pos"] - q["pos"]).max() < 1e-10
        assert np.abs(out["vel"] - q["vel"]).max() < 1e-9
        gpu.build_tree(out, 16, 256)
        rms, p99 = subset_errors(gpu, out, fp, 32)
        assert rms < 1e-3 and p99 < 5e-3, (rms, p99)
    finally:
        gpu.set_precision(bl.BH_FP32)
