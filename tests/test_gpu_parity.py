"""GPU parity through the C ABI against the oracle and the reference fixtures.

Bars: keys, permutation, sorted particles, node arrays, centres, second
moments and group roots bit-exact (reference node layout); per-particle
interaction-list lengths exact (so the MAC decisions are identical);
accelerations within 1e-12 (FP64 walk) and 1e-5 (FP32 walk, batch sums
accumulated in FP64), global-normalised, against the FP64 oracle.
"""
import numpy as np
import pytest

import golden_io
from bhgpu import ic, layout
from bhgpu import lib as bl

pytestmark = pytest.mark.gpu

TREE_FIELDS = ("keys", "perm", "particles", "nodes", "center", "iab", "groups")
TOL = {bl.BH_FP64: 1e-12, bl.BH_FP32: 1e-5}


def rel_err(a, b):
    den = (b ** 2).sum()
    return float(np.sqrt(((a - b) ** 2).sum() / den)) if den > 0 else float(np.abs(a - b).max())


def check_tree(gpu, orc, p, ml, gs):
    print(f"tree n={len(p)} max_leaf={ml} group_size={gs}", flush=True)
    gpu.build_tree(p, ml, gs)
    ge = gpu.export()
    oe = orc.build(p, ml, gs).export()
    for k in TREE_FIELDS:
        assert golden_io.bytes_equal(ge[k], oe[k]), k
    assert np.array_equal(ge["domain"], oe["domain"])
    return ge


CASES = [(1, 8, 64, 1), (2, 8, 64, 1), (9, 8, 64, 2), (200, 4, 16, 1), (2000, 8, 64, 4), (4096, 16, 256, 3),
         (4097, 16, 256, 3), (20000, 16, 64, 5), (20000, 16, 8192, 5), (5000, 1, 64, 6), (3000, 0, 64, 7),
         (70000, 32, 256, 9),
         # leaf capacities whose windows (max_leaf + 1) need 3, 6, 6 and 8
         # doubling levels in the count kernel's window table (7: exactly 8)
         (30000, 7, 64, 10), (30000, 63, 256, 11), (40000, 100, 512, 12), (40000, 300, 1024, 13)]


@pytest.mark.parametrize("n,ml,gs,seed", CASES)
def test_tree_bit_exact(gpu, orc, n, ml, gs, seed):
    check_tree(gpu, orc, ic.random_box(n, seed, 10.0), ml, gs)


def test_tree_plummer_and_uniform(gpu, orc):
    check_tree(gpu, orc, ic.plummer(65536, 1), 16, 256)
    check_tree(gpu, orc, ic.uniform_cube(50000, 2), 8, 256)


def test_tree_leaves_beyond_one_staging_chunk(gpu, orc):
    """Leaves whose particles span several of the upward pass's 1,024-particle
    LDS chunks: a 2,500-fold duplicate (one depth-21 leaf) and max_leaf
    1,024, the largest allowed; moments bit-exact."""
    p = ic.random_box(6000, 13, 10.0)
    p["pos"][100:2600] = p["pos"][5]
    check_tree(gpu, orc, p, 1024, 4096)
    check_tree(gpu, orc, ic.plummer(30000, 14), 1024, 8192)


def test_tree_duplicates_and_degenerate(gpu, orc):
    p = ic.random_box(3000, 3, 10.0)
    p["pos"][:1000] = p["pos"][17]
    check_tree(gpu, orc, p, 4, 16)
    q = ic.random_box(50, 3, 10.0)
    q["pos"][:] = q["pos"][0]
    check_tree(gpu, orc, q, 4, 16)
    r = ic.random_box(500, 4, 10.0)
    r["pos"][:, 2] = 1.5  # flat in z
    check_tree(gpu, orc, r, 8, 64)


@pytest.mark.parametrize("name", golden_io.NAMES)
def test_reference_fixtures(gpu, name):
    g = golden_io.load(name)
    print(f"fixture {name}", flush=True)
    gpu.build_tree(g["particles"], g["max_leaf"], g["group_size"])
    ge = gpu.export()
    for k in ("keys", "perm", "nodes", "center", "iab", "groups"):
        assert golden_io.bytes_equal(ge[k], g[k]), k
    for prec in (bl.BH_FP64, bl.BH_FP32):
        gpu.set_precision(prec)
        acc, npp, npc, _ = gpu.accel(g["fp"], g["target_size"])
        assert np.array_equal(npp, g["n_pp"]) and np.array_equal(npc, g["n_pc"])
        assert rel_err(acc, g["acc"]) < TOL[prec]
    gpu.set_precision(bl.BH_FP32)


WALKS = [(2000, 8, 64, 32, 0.5, 1e-6, True), (2000, 8, 64, 32, 0.5, 1e-6, False), (2000, 8, 64, 32, 0.5, 0.0, True),
         (20000, 16, 64, 32, 0.5, 1e-6, True), (20000, 16, 8192, 32, 0.6, 1e-6, True), (5000, 1, 64, 32, 0.5, 1e-6, True),
         (5000, 4, 256, 100, 0.5, 1e-6, True), (3000, 8, 64, 0, 0.4, 1e-6, True), (3000, 2, 512, 300, 0.7, 1e-4, True),
         (200, 4, 16, 32, 0.0, 1e-6, True)]


@pytest.mark.parametrize("n,ml,gs,ts,theta,eps2,quad", WALKS)
def test_walk_parity(gpu, orc, n, ml, gs, ts, theta, eps2, quad):
    p = ic.random_box(n, 11, 10.0)
    fp = layout.force_params(theta, eps2, 1.0, quad)
    ot = orc.build(p, ml, gs)
    oa, opp, opc, otot = ot.grouped_pass(fp, ts)
    gpu.build_tree(p, ml, gs)
    for prec in (bl.BH_FP64, bl.BH_FP32):
        print(f"walk n={n} ml={ml} gs={gs} ts={ts} theta={theta} eps2={eps2} quad={quad} prec={prec}", flush=True)
        gpu.set_precision(prec)
        ga, gpp, gpc, gtot = gpu.accel(fp, ts)
        assert np.array_equal(gpp, opp) and np.array_equal(gpc, opc) and np.array_equal(gtot, otot)
        assert rel_err(ga, oa) < TOL[prec]
    gpu.set_precision(bl.BH_FP32)


@pytest.mark.parametrize("mode", [bl.BH_WALK_GROUPED, bl.BH_WALK_PER_TARGET])
@pytest.mark.parametrize("ts,ranges", [(64, ((1, 32), (33, 42), (43, 53), (54, 64))),
                                       (128, ((1, 32), (33, 64), (65, 96), (97, 128)))])
def test_wide_targets(gpu, orc, mode, ts, ranges):
    """Targets of 33-128 particles go through the group evaluator's layouts
    (two slots for 33-42, 32 + the rest for 43-53, one slot for 54-64, 64 +
    the rest for 65-96, two full slots for 97-128); smaller ones through the
    target evaluator. Target sizes 64 and 128 over leaves of 16 give all of
    them; lists exact, sums within tolerance, in both walk schedules."""
    p = ic.plummer(40000, 23)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    ot = orc.build(p, 16, 256)
    sizes = ot.export()["nodes"]["particle_count"][ot.targets(ts)]
    for lo, hi in ranges:
        assert ((sizes >= lo) & (sizes <= hi)).any(), (lo, hi)
    oa, opp, opc, otot = ot.grouped_pass(fp, ts)
    gpu.set_walk_mode(mode)
    try:
        gpu.build_tree(p, 16, 256)
        for prec in (bl.BH_FP64, bl.BH_FP32):
            print(f"wide targets ts={ts} mode={mode} prec={prec}", flush=True)
            gpu.set_precision(prec)
            ga, gpp, gpc, gtot = gpu.accel(fp, ts)
            assert np.array_equal(gpp, opp) and np.array_equal(gpc, opc) and np.array_equal(gtot, otot)
            assert rel_err(ga, oa) < TOL[prec]
    finally:
        gpu.set_walk_mode(bl.BH_WALK_GROUPED)
        gpu.set_precision(bl.BH_FP32)


@pytest.mark.parametrize("mode", [bl.BH_WALK_GROUPED, bl.BH_WALK_PER_TARGET])
@pytest.mark.parametrize("theta,kind", [(0.1, "plummer"), (0.5, "plummer"), (0.8, "hernquist"), (0.5, "hernquist")])
def test_walk_modes_clustered(gpu, orc, mode, theta, kind):
    # theta 0.1 defers more nodes than a group may hold: its groups fall back
    # to walks from the root
    p = getattr(ic, kind)(30000, 21)
    fp = layout.force_params(theta, 1e-4, 1.0, True)
    ot = orc.build(p, 8, 256)
    oa, opp, opc, otot = ot.grouped_pass(fp, 32)
    gpu.set_walk_mode(mode)
    try:
        gpu.build_tree(p, 8, 256)
        for prec in (bl.BH_FP64, bl.BH_FP32):
            print(f"walk mode={mode} {kind} theta={theta} prec={prec}", flush=True)
            gpu.set_precision(prec)
            ga, gpp, gpc, gtot = gpu.accel(fp, 32)
            assert np.array_equal(gpp, opp) and np.array_equal(gpc, opc) and np.array_equal(gtot, otot)
            assert rel_err(ga, oa) < TOL[prec]
    finally:
        gpu.set_walk_mode(bl.BH_WALK_GROUPED)
        gpu.set_precision(bl.BH_FP32)


@pytest.mark.parametrize("theta,quad", [(0.5, True), (0.7, True), (0.5, False)])
def test_per_node_cross_check(gpu, orc, theta, quad):
    """target_size 0 makes every leaf a target: the lists of the reference's
    per-node path (traversal_ttg.h:156-196), restated in the oracle."""
    p = ic.hernquist(20000, 9)
    fp = layout.force_params(theta, 1e-4, 1.0, quad)
    oa, opp, opc = orc.build(p, 8, 64).per_node_pass(fp)
    if not quad:
        opp, opc = opp + opc, opc * 0
    gpu.build_tree(p, 8, 64)
    for prec in (bl.BH_FP64, bl.BH_FP32):
        print(f"per-node cross-check theta={theta} quad={quad} prec={prec}", flush=True)
        gpu.set_precision(prec)
        ga, gpp, gpc, _ = gpu.accel(fp, 0)
        assert np.array_equal(gpp, opp) and np.array_equal(gpc, opc)
        assert rel_err(ga, oa) < TOL[prec]
    gpu.set_precision(bl.BH_FP32)


def test_large_targets_span_chunks(gpu, orc):
    # targets of more than 64 particles (target_size 200 and a 150-fold duplicate leaf)
    p = ic.random_box(4000, 12, 10.0)
    p["pos"][:150] = p["pos"][0]
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    ot = orc.build(p, 4, 512)
    oa, opp, opc, _ = ot.grouped_pass(fp, 200)
    gpu.build_tree(p, 4, 512)
    for prec in (bl.BH_FP64, bl.BH_FP32):
        gpu.set_precision(prec)
        ga, gpp, gpc, _ = gpu.accel(fp, 200)
        assert np.array_equal(gpp, opp) and np.array_equal(gpc, opc)
        assert rel_err(ga, oa) < TOL[prec]
    gpu.set_precision(bl.BH_FP32)


def packed_runs(export, tsm, cap=128):
    """Number of walk units the walk packs from runs (two or more) of
    consecutive sibling groups that are single targets, under nodes with
    more than max(tsm, cap) particles, greedily up to cap particles."""
    nodes = export["nodes"]
    leaf = nodes["leaf"].astype(bool)
    cnt = nodes["particle_count"].astype(np.int64)
    fc = nodes["first_child"].astype(np.int64)
    nch = nodes["num_children"].astype(np.int64)
    units = 0
    for i in np.nonzero(~leaf & (cnt > max(tsm, cap)))[0]:
        run_len, run_n = 0, 0
        for c in list(range(fc[i], fc[i] + nch[i])) + [None]:
            single = c is not None and (leaf[c] or cnt[c] <= tsm) and cnt[c] <= cap
            if single and run_len > 0 and run_n + cnt[c] <= cap:
                run_len, run_n = run_len + 1, run_n + cnt[c]
                continue
            units += run_len > 1
            run_len, run_n = (1, cnt[c]) if single else (0, 0)
    return units


def walk_group_sizes(export, tsm, cap=128):
    """Particle counts of the walk groups that run a shared phase: topmost
    nodes with leaf || count <= max(tsm, cap) that are not targets."""
    nodes = export["nodes"]
    leaf = nodes["leaf"].astype(bool)
    cnt = nodes["particle_count"].astype(np.int64)
    fc = nodes["first_child"].astype(np.int64)
    nch = nodes["num_children"].astype(np.int64)
    par = np.full(len(cnt), -1)
    for i in np.nonzero(~leaf)[0]:
        par[fc[i]:fc[i] + nch[i]] = i
    pred = leaf | (cnt <= max(tsm, cap))
    top = pred & ((par < 0) | ~pred[np.maximum(par, 0)])
    return cnt[top & ~leaf & (cnt > tsm)]


@pytest.mark.parametrize("kind,n,ml", [("plummer", 60000, 32), ("hernquist", 40000, 16), ("plummer", 30000, 8)])
def test_group_layouts_by_size(gpu, orc, kind, n, ml):
    """The shared phase lays a unit's particles over the lanes by its size
    (two per lane in slices up to 42, 32 in two slices plus a sliced rest
    up to 53, one per lane up to 64, a sliced tail up to 96, two per lane up
    to 128), and runs of sibling single-target
    groups are packed into units: every range and packed units occur here,
    and lists and forces match the oracle."""
    p = getattr(ic, kind)(n, 13)
    fp = layout.force_params(0.5, 1e-6, 1.0, True)
    ot = orc.build(p, ml, 256)
    ex = ot.export()
    sizes = walk_group_sizes(ex, 32)
    for lo, hi in ((33, 42), (43, 53), (54, 64), (65, 96), (97, 128)):
        assert ((sizes >= lo) & (sizes <= hi)).any(), (lo, hi)
    assert packed_runs(ex, 32) > 0
    oa, opp, opc, otot = ot.grouped_pass(fp, 32)
    gpu.build_tree(p, ml, 256)
    for prec in (bl.BH_FP64, bl.BH_FP32):
        print(f"group layouts {kind} n={n} ml={ml} prec={prec}", flush=True)
        gpu.set_precision(prec)
        ga, gpp, gpc, gtot = gpu.accel(fp, 32)
        assert np.array_equal(gpp, opp) and np.array_equal(gpc, opc) and np.array_equal(gtot, otot)
        assert rel_err(ga, oa) < TOL[prec]
    gpu.set_precision(bl.BH_FP32)


def test_theta0_equals_direct_sum(gpu):
    p = ic.random_box(1500, 1, 10.0)
    fp = layout.force_params(0.0, 1e-6, 1.0, True)
    gpu.build_tree(p, 4, 16)
    gpu.set_precision(bl.BH_FP64)
    acc = gpu.accel(fp, 0)[0]
    gpu.set_precision(bl.BH_FP32)
    perm = gpu.export_perm()
    direct = gpu.direct_sum(p, perm, 1e-6)
    assert rel_err(acc, direct) < 1e-12


def test_direct_sum_matches_oracle(gpu, orc):
    p = ic.random_box(3000, 9, 10.0)
    idx = np.arange(0, 3000, 7, dtype=np.uint32)
    assert rel_err(gpu.direct_sum(p, idx, 1e-6), orc.brute_subset(p, idx, 1e-6)) < 1e-14
    assert rel_err(gpu.direct_sum(p[:500], None, 0.0, 2.0), orc.brute_force(p[:500], 0.0, 2.0)) < 1e-14


@pytest.mark.parametrize("prec", [bl.BH_FP64, bl.BH_FP32])
def test_step_parity(gpu, orc, prec):
    p = ic.plummer(20000, 1)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    gpu.set_precision(prec)
    out, t = gpu.step(p, fp, 16, 256, 32, 5e-4, 1e-3)
    gpu.set_precision(bl.BH_FP32)
    ref, _ = orc.step(p, fp, 16, 256, 32, 5e-4, 1e-3)
    assert np.array_equal(out["id"], ref["id"]) and np.array_equal(out["mass"], ref["mass"])
    tol = 1e-12 if prec == bl.BH_FP64 else 1e-5
    dv = out["vel"] - ref["vel"]
    assert np.sqrt((dv ** 2).sum() / (ref["vel"] ** 2).sum()) < tol
    assert np.abs(out["pos"] - ref["pos"]).max() < 1e-9
    assert t.tree_build_ms > 0 and t.traversal_ms > 0 and t.exchange_ms == 0


def test_step_no_drift_keeps_positions(gpu):
    p = ic.random_box(5000, 3, 5.0, vel_spread=0.2)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    out, _ = gpu.step(p, fp, 8, 256, 32, 1e-3, 0.0)
    order = np.argsort(out["id"])
    assert np.array_equal(out["pos"][order], p["pos"])


def test_leapfrog_resident_state_matches_oracle(gpu, orc):
    # the leapfrog schedule of main.cc:199-246 on resident state, 10 steps
    p = ic.plummer(8192, 2)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    dt, steps = 1e-3, 10
    gpu.set_precision(bl.BH_FP64)
    gpu.state_upload(p)
    q = p
    sched = [(dt / 2, dt)] + [(dt, dt)] * (steps - 1) + [(dt / 2, 0.0)]
    for k, d in sched:
        gpu.state_step(fp, 16, 256, 32, k, d)
        q, _ = orc.step(q, fp, 16, 256, 32, k, d)
    out = gpu.state_download()
    gpu.set_precision(bl.BH_FP32)
    assert np.array_equal(out["id"], q["id"])
    assert np.abs(out["pos"] - q["pos"]).max() < 1e-10
    assert np.abs(out["vel"] - q["vel"]).max() < 1e-9


def test_energy_conservation_gpu(gpu, orc):
    # test_energy_conservation.cc: cold cube, 30 steps, |dE/E| <= 5%
    p = ic.random_box(1000, 42, 5.0)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    dt = 1e-3
    e0 = orc.kinetic_energy(p) + orc.potential_energy(p)
    gpu.state_upload(p)
    gpu.state_step(fp, 8, 256, 32, dt / 2, dt)
    for _ in range(29):
        gpu.state_step(fp, 8, 256, 32, dt, dt)
    gpu.state_step(fp, 8, 256, 32, dt / 2, 0.0)
    q = gpu.state_download()
    e1 = orc.kinetic_energy(q) + orc.potential_energy(q)
    assert abs((e1 - e0) / e0) <= 0.05


def test_error_paths(gpu):
    p = ic.random_box(1000, 1, 1.0)
    fp = layout.force_params(0.5, 1e-6, 1.0, True)
    bad = p.copy()
    bad["pos"][10, 0] = np.inf
    with pytest.raises(bl.BHError) as e:
        gpu.step(bad, fp, 8, 256, 32, 1e-3, 1e-3)
    assert e.value.code == 2
    bad = p.copy()
    bad["mass"][3] = np.nan
    with pytest.raises(bl.BHError) as e:
        gpu.build_tree(bad, 8, 64)
    assert e.value.code == 2
    with pytest.raises(bl.BHError) as e:
        gpu.step(p, fp, 8, 0, 32, 1e-3, 1e-3)
    assert e.value.code == 1
    with pytest.raises(bl.BHError) as e:
        gpu.step(p, fp, 5000, 256, 32, 1e-3, 1e-3)
    assert e.value.code == 6
    with pytest.raises(bl.BHError) as e:
        gpu.step(p, layout.force_params(-1.0, 0.0, 1.0, True), 8, 256, 32, 1e-3, 1e-3)
    assert e.value.code == 1
    out, _ = gpu.step(p[:0], fp, 8, 256, 32, 1e-3, 1e-3)
    assert len(out) == 0
    # the context is still usable afterwards
    out, _ = gpu.step(p, fp, 8, 256, 32, 1e-3, 1e-3)
    assert np.array_equal(np.sort(out["id"]), np.arange(1000))


def test_deterministic_replay(gpu):
    p = ic.plummer(30000, 5)
    fp = layout.force_params(0.5, 1e-6, 1.0, True)
    a, _ = gpu.step(p, fp, 16, 256, 32, 1e-3, 1e-3)
    b, _ = gpu.step(p, fp, 16, 256, 32, 1e-3, 1e-3)
    assert golden_io.bytes_equal(a, b)


@pytest.mark.parametrize("tile", [16, 32, 48, 64, 96, 128])
def test_walk_tile_keeps_lists(gpu, orc, tile):
    """The walk tile (particles per walk unit) changes how targets share
    walks, never what a target interacts with: lists exact, forces within
    the precision's bar, for every tile."""
    p = ic.plummer(30000, 91)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    oa, opp, opc, otot = orc.build(p, 8, 256).grouped_pass(fp, 32)
    gpu.set_walk_tile(tile)
    try:
        gpu.build_tree(p, 8, 256)
        for prec in (bl.BH_FP64, bl.BH_FP32):
            print(f"walk tile {tile} prec={prec}", flush=True)
            gpu.set_precision(prec)
            ga, gpp, gpc, gtot = gpu.accel(fp, 32)
            assert np.array_equal(gpp, opp) and np.array_equal(gpc, opc) and np.array_equal(gtot, otot)
            assert rel_err(ga, oa) < TOL[prec]
    finally:
        gpu.set_walk_tile(128)
        gpu.set_precision(bl.BH_FP32)
    with pytest.raises(bl.BHError):
        gpu.set_walk_tile(8)


def test_incomplete_walk_reported(gpu):
    """The reference throws when a walk leaves interactions missing
    (integrate_ttg.h:104-106). With the walk's stack capped below the tree's
    depth (a diagnostic limit: the kernel still finishes normally) the walk
    meets nodes it may not open, flags it, and the step and bh_accel return
    BH_ERR_INCOMPLETE; with the full stack the same input steps normally."""
    p = ic.plummer(20000, 93)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    try:
        gpu.debug_walk_stack_limit(3)
        with pytest.raises(bl.BHError) as e:
            gpu.step(p, fp, 16, 256, 32, 1e-3, 1e-3)
        assert e.value.code == 3 and "incomplete" in str(e.value)
        gpu.build_tree(p, 16, 256)
        with pytest.raises(bl.BHError) as e:
            gpu.accel(fp, 32)
        assert e.value.code == 3
    finally:
        gpu.debug_walk_stack_limit(22)
    out, _ = gpu.step(p, fp, 16, 256, 32, 1e-3, 1e-3)
    assert np.array_equal(np.sort(out["id"]), np.arange(len(p)))
    with pytest.raises(bl.BHError):
        gpu.debug_walk_stack_limit(0)


@pytest.mark.parametrize("name", golden_io.DIGEST_NAMES)
def test_reference_digests_at_depth(gpu, name):
    """The GPU tree (reference node numbering, after the host renumbering of
    several backbone levels) and the walk's per-particle list lengths hash to
    the digests of the reference-compiled outputs at 131,072 particles."""
    d = golden_io.load_digest(name)
    p = golden_io.digest_input(d)
    gpu.build_tree(p, d["max_leaf"], d["group_size"])
    e = gpu.export()
    got = {k: golden_io.sha(e[k]) for k in ("keys", "perm", "nodes", "center", "iab", "groups")}
    _, npp, npc, _ = gpu.accel(layout.force_params(d["theta"], d["eps2"], 1.0, True), d["target_size"])
    got |= {"n_pp": golden_io.sha(npp), "n_pc": golden_io.sha(npc)}
    assert got == d["sha256"]
