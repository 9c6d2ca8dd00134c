"""Pins the oracle (oracle/bh_oracle.cpp) on the CPU.

Against the reference's own compiled functions (oracle/_ref, where built),
against the committed reference fixtures, and against what the mathematics
fixes: the two-body force, theta = 0 = brute force, closed-form multipoles,
the unique (key, index) sort order, tree invariants, energy conservation.
Tests mirror tests/test_force_kernel.cc, tests/test_bh_vs_bruteforce.cc and
tests/test_energy_conservation.cc of the reference.
"""
import numpy as np
import pytest

import golden_io
from bhgpu import ic, layout


def rel_err(a, b):
    """Global-normalised error, test_bh_vs_bruteforce.cc:33-40."""
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


# ---------------------------------------------------------------- force kernel
def test_two_body_kat(orc):
    # test_force_kernel.cc:12-25: G=2, m=3, d=4 -> a = (0.375, 0, 0)
    a = orc.accel_from_source([0, 0, 0], 3.0, [4.0, 0, 0], np.zeros(6), 0.0, False, 2.0)
    assert a[0] == 0.375 and a[1] == 0.0 and a[2] == 0.0


def test_softening_weakens_and_zero_iab(orc):
    a = orc.accel_from_source([0, 0, 0], 3.0, [4.0, 0, 0], np.zeros(6), 0.0, False, 2.0)
    s = orc.accel_from_source([0, 0, 0], 3.0, [4.0, 0, 0], np.zeros(6), 4.0, False, 2.0)
    q = orc.accel_from_source([0, 0, 0], 3.0, [4.0, 0, 0], np.zeros(6), 0.0, True, 2.0)
    assert np.linalg.norm(s) < np.linalg.norm(a)
    assert np.array_equal(a, q)


def test_quadrupole_matches_expansion(orc):
    # A point pair (masses 1 at com +- h e_x) seen from far away along y: the
    # exact field and the monopole+quadrupole expansion agree to O((h/r)^4).
    h, r = 0.01, 1.0
    tp = np.array([0.0, r, 0.0])
    exact = orc.accel_from_source(tp, 1.0, [h, 0, 0], np.zeros(6), 0.0, False, 1.0) + \
        orc.accel_from_source(tp, 1.0, [-h, 0, 0], np.zeros(6), 0.0, False, 1.0)
    iab = np.array([2 * h * h, 0, 0, 0, 0, 0])
    quad = orc.accel_from_source(tp, 2.0, [0, 0, 0], iab, 0.0, True, 1.0)
    mono = orc.accel_from_source(tp, 2.0, [0, 0, 0], np.zeros(6), 0.0, False, 1.0)
    assert np.linalg.norm(quad - exact) < 1e-3 * np.linalg.norm(mono - exact)


def test_force_kernel_matches_reference(orc, ref):
    rng = np.random.default_rng(0)
    for _ in range(200):
        tp, com = rng.normal(size=3), rng.normal(size=3) * 3
        iab = rng.normal(size=6)
        m, eps2, G = rng.uniform(0.1, 2), rng.choice([0.0, 1e-3]), rng.uniform(0.5, 2)
        for quad in (False, True):
            assert np.array_equal(orc.accel_from_source(tp, m, com, iab, eps2, quad, G),
                                  ref.accel_from_source(tp, m, com, iab, eps2, quad, G))


# ------------------------------------------------------------- domain and keys
def test_spread_bits_kat(orc):
    assert orc.spread_bits3(0) == 0
    assert orc.spread_bits3(1) == 1
    assert orc.spread_bits3(0b11) == 0b1001
    assert orc.spread_bits3((1 << 21) - 1) == 0x1249249249249249
    for v in (5, 12345, 1 << 20, 0x155555):
        expect = sum(((v >> b) & 1) << (3 * b) for b in range(21))
        assert orc.spread_bits3(v) == expect


def test_morton_key_kats(orc):
    p = layout.make_particles(np.array([[-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]]))
    dom = orc.domain(p)
    assert np.array_equal(dom[:3], [0.0, 0.0, 0.0])
    assert dom[3] == 1.0 * (1.0 + 1e-9)
    keys = orc.morton_keys(p, dom)
    assert keys[0] == 0  # low corner
    assert keys[1] == (1 << 63) - 1  # high corner: all 21 digits 7
    # the centre is the first cell of the upper octant on every axis
    c = orc.morton_keys(layout.make_particles(np.zeros((1, 3))), dom)[0]
    assert c == 7 << 60
    # x is the low bit of a digit, z the high bit
    px = orc.morton_keys(layout.make_particles(np.array([[0.5, -0.5, -0.5]])), dom)[0]
    pz = orc.morton_keys(layout.make_particles(np.array([[-0.5, -0.5, 0.5]])), dom)[0]
    assert px >> 60 == 1 and pz >> 60 == 4


def test_degenerate_domain(orc):
    p = layout.make_particles(np.full((3, 3), 2.5))
    dom = orc.domain(p)
    assert np.array_equal(dom, [2.5, 2.5, 2.5, 1.0 + 1e-9])


def test_keys_match_reference(orc, ref):
    for p in (ic.random_box(5000, 1, 10.0), ic.plummer(5000, 2), ic.uniform_cube(5000, 3)):
        dom = orc.domain(p)
        assert np.array_equal(dom, ref.domain(p))
        assert np.array_equal(orc.morton_keys(p, dom), ref.morton_keys(p, dom))


# ---------------------------------------------------------------- sort + tree
@pytest.mark.parametrize("n,seed", [(1, 1), (7, 2), (5000, 3), (20000, 4)])
def test_sort_is_unique_key_index_order(orc, n, seed):
    p = ic.random_box(n, seed, 10.0)
    p["pos"][: n // 3] = p["pos"][0]  # ties broken by input index
    e = orc.build(p, 8, 64).export()
    keys = orc.morton_keys(p, e["domain"])
    order = np.lexsort((np.arange(n), keys))
    assert np.array_equal(e["perm"], order)
    assert np.array_equal(e["keys"], keys[order])
    assert np.array_equal(e["particles"], p[order])


def tree_invariants(e, n, max_leaf):
    nodes = e["nodes"]
    leaf = nodes["leaf"].astype(bool)
    cnt = nodes["particle_count"]
    assert cnt[0] == n and nodes["particle_begin"][0] == 0
    cover = np.zeros(n, np.int64)
    for nd in nodes[leaf]:
        cover[nd["particle_begin"]: nd["particle_begin"] + nd["particle_count"]] += 1
    assert np.all(cover == 1)  # every particle in exactly one leaf
    assert np.all((cnt[leaf] <= max_leaf) | (nodes["depth"][leaf] >= 21))
    assert np.all(cnt[~leaf] > max_leaf) and np.all(nodes["depth"] <= 21)
    for i in np.nonzero(~leaf)[0]:
        nd = nodes[i]
        ch = nodes[nd["first_child"]: nd["first_child"] + nd["num_children"]]
        assert ch["particle_count"].sum() == nd["particle_count"]
        assert np.all(np.diff(ch["particle_begin"].astype(np.int64)) > 0)
        assert np.all(ch["depth"] == nd["depth"] + 1)
        assert ch["particle_begin"][0] == nd["particle_begin"]


@pytest.mark.parametrize("n,ml,gs,seed", [(200, 4, 16, 1), (2000, 8, 64, 4), (20000, 16, 64, 5),
                                          (20000, 16, 8192, 5), (5000, 1, 64, 6), (3000, 0, 64, 7)])
def test_tree_matches_reference(orc, ref, n, ml, gs, seed):
    p = ic.random_box(n, seed, 10.0)
    o = orc.build(p, ml, gs).export()
    r = ref.build(p, ml, gs).export()
    for k in ("keys", "perm", "nodes", "center", "iab", "groups"):
        assert golden_io.bytes_equal(o[k], r[k]), k
    tree_invariants(o, n, ml)


def test_multipoles_closed_form(orc):
    p = ic.plummer(3000, 5)
    e = orc.build(p, 8, 64).export()
    sp, nodes = e["particles"], e["nodes"]
    for i in range(0, len(nodes), 37):
        nd = nodes[i]
        s = sp[nd["particle_begin"]: nd["particle_begin"] + nd["particle_count"]]
        m = s["mass"]
        M = m.sum()
        com = (m[:, None] * s["pos"]).sum(0) / M
        d = s["pos"] - com
        full = np.einsum("i,ij,ik->jk", m, d, d)
        want = np.array([full[0, 0], full[1, 1], full[2, 2], full[0, 1], full[0, 2], full[1, 2]])
        assert abs(nd["mass"] - M) <= 1e-12 * M
        assert np.allclose(nd["com"], com, rtol=0, atol=1e-12 * (1 + np.abs(com).max()))
        assert np.allclose(e["iab"][i], want, rtol=1e-9, atol=1e-12 * M)


def test_duplicates_stop_at_depth_21(orc):
    p = ic.random_box(100, 3, 10.0)
    p["pos"][:40] = p["pos"][0]
    e = orc.build(p, 4, 16).export()
    tree_invariants(e, 100, 4)
    big = e["nodes"][e["nodes"]["particle_count"] >= 40]
    assert big["depth"].max() == 21 and big[big["depth"] == 21]["leaf"].all()


def test_golden_fixtures(orc):
    for name in golden_io.NAMES:
        g = golden_io.load(name)
        t = orc.build(g["particles"], g["max_leaf"], g["group_size"])
        e = t.export()
        for k in ("keys", "perm", "nodes", "center", "iab", "groups"):
            assert golden_io.bytes_equal(e[k], g[k]), (name, k)
        assert np.array_equal(e["domain"], g["domain"])
        acc, npp, npc, _ = t.grouped_pass(g["fp"], g["target_size"])
        assert np.array_equal(npp, g["n_pp"]) and np.array_equal(npc, g["n_pc"]), name
        assert rel_err(acc, g["acc"]) < 1e-13, name


# ----------------------------------------------------------------------- walk
def test_theta0_equals_brute_force(orc):
    # test_bh_vs_bruteforce.cc:52-76
    p = ic.random_box(200, 1, 10.0)
    fp = layout.force_params(0.0, 1e-6, 1.0, True)
    for gs in (16, 1 << 20):
        t = orc.build(p, 4, gs)
        e = t.export()
        acc = t.grouped_pass(fp, 0)[0]
        assert rel_err(acc, orc.brute_force(e["particles"], 1e-6)) < 1e-8


def test_brute_force_matches_reference(orc, ref):
    p = ic.random_box(500, 2, 10.0)
    assert np.array_equal(orc.brute_force(p, 1e-6), ref.brute_force(p, 1e-6))


def test_quadrupole_beats_monopole(orc):
    # test_bh_vs_bruteforce.cc:78-113: two clusters at +-50
    a = ic.random_box(100, 2, 1.0, center=(-50, 0, 0))
    b = ic.random_box(100, 3, 1.0, center=(50, 0, 0))
    p = np.concatenate([a, b])
    p["id"] = np.arange(200)
    errs = {}
    for quad in (False, True):
        fp = layout.force_params(0.6, 1e-6, 1.0, quad)
        for gs in (16, 1 << 20):
            t = orc.build(p, 4, gs)
            e = t.export()
            errs[(quad, gs)] = rel_err(t.grouped_pass(fp, 0)[0], orc.brute_force(e["particles"], 1e-6))
    for gs in (16, 1 << 20):
        assert errs[(True, gs)] < errs[(False, gs)]


@pytest.mark.parametrize("n,ml,gs,ts,quad,eps2", [(2000, 8, 64, 32, True, 1e-6), (2000, 1, 64, 32, False, 1e-6),
                                                  (20000, 16, 64, 32, True, 1e-6), (2000, 8, 64, 32, True, 0.0),
                                                  (20000, 16, 8192, 32, True, 1e-6)])
def test_grouped_walk_equals_target_walks(orc, ref, n, ml, gs, ts, quad, eps2):
    # grouped_ttg.h:42-51: the shared walk yields each target's own root walk
    p = ic.random_box(n, 5, 10.0)
    fp = layout.force_params(0.5, eps2, 1.0, quad)
    t = orc.build(p, ml, gs)
    ga, gpp, gpc, gt = t.grouped_pass(fp, ts)
    ta, tpp, tpc, tt = t.target_walk(fp, ts)
    assert np.array_equal(gpp, tpp) and np.array_equal(gpc, tpc) and np.array_equal(gt, tt)
    assert rel_err(ga, ta) < 1e-13
    ra, rpp, rpc = ref.build(p, ml, gs).target_walk(0.5, eps2, 1.0, quad, ts)
    assert np.array_equal(gpp, rpp) and np.array_equal(gpc, rpc)
    assert rel_err(ga, ra) < 1e-13


def test_grouped_matches_per_leaf_walk(orc):
    # test_bh_vs_bruteforce.cc:116-138: target_size <= max_leaf = per-leaf walks
    p = ic.random_box(2000, 4, 10.0)
    fp = layout.force_params(0.5, 1e-6, 1.0, True)
    t = orc.build(p, 8, 64)
    a1 = t.grouped_pass(fp, 8)[0]
    a2 = t.target_walk(fp, 1)[0]  # targets = leaves either way
    assert rel_err(a1, a2) < 1e-8


def test_shared_targets_not_less_accurate(orc):
    # test_bh_vs_bruteforce.cc:140-166
    p = ic.random_box(2000, 6, 10.0)
    fp0 = layout.force_params(0.0, 1e-6, 1.0, False)
    t = orc.build(p, 1, 64)
    e = t.export()
    bf = orc.brute_force(e["particles"], 1e-6)
    assert rel_err(t.grouped_pass(fp0, 32)[0], bf) < 1e-8
    fp = layout.force_params(0.5, 1e-6, 1.0, False)
    assert rel_err(t.grouped_pass(fp, 32)[0], bf) <= rel_err(t.grouped_pass(fp, 1)[0], bf)


def test_multi_subtree_build(orc):
    # test_bh_vs_bruteforce.cc:168-198: backbone + subtrees, groups inside / above them
    p = ic.random_box(20000, 5, 10.0)
    fp = layout.force_params(0.5, 1e-6, 1.0, True)
    base = orc.build(p, 16, 64).grouped_pass(fp, 32)[0]
    for gs in (64, 8192):
        t = orc.build(p, 16, gs)
        assert rel_err(t.grouped_pass(fp, 32)[0], base) < 1e-8


# ------------------------------------------------------------------ integrator
def test_step_is_build_walk_kick_drift(orc):
    p = ic.random_box(3000, 8, 5.0, vel_spread=0.1)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    out, _ = orc.step(p, fp, 8, 256, 32, 5e-4, 1e-3)
    t = orc.build(p, 8, 256)
    e = t.export()
    acc = t.grouped_pass(fp, 32)[0]
    vel = e["particles"]["vel"] + acc * 5e-4
    pos = e["particles"]["pos"] + vel * 1e-3
    assert np.array_equal(out["id"], e["particles"]["id"])
    assert np.array_equal(out["vel"], vel) and np.array_equal(out["pos"], pos)


def test_kick_drift_matches_reference(orc, ref):
    p = ic.random_box(1000, 9, 5.0, vel_spread=0.3)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    for drift in (1e-3, 0.0):
        out, _ = orc.step(p, fp, 8, 256, 32, 5e-4, drift)
        t = orc.build(p, 8, 256)
        e = t.export()
        r = ref.kick_drift(e["particles"], t.grouped_pass(fp, 32)[0], 5e-4, drift)
        assert golden_io.bytes_equal(out, r)


def test_energies_match_reference(orc, ref):
    p = ic.random_box(800, 10, 5.0, vel_spread=0.5)
    assert orc.kinetic_energy(p) == ref.kinetic_energy(p)
    assert abs(orc.potential_energy(p) - ref.potential_energy(p)) <= 1e-12 * abs(ref.potential_energy(p))


def test_energy_conservation(orc):
    # test_energy_conservation.cc:13-71 (cold cube, 30 steps, drift <= 5%),
    # with the leapfrog schedule of main.cc:199-246
    p = ic.random_box(1000, 42, 5.0)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    dt, steps = 1e-3, 30
    e0 = orc.kinetic_energy(p) + orc.potential_energy(p)
    p, _ = orc.step(p, fp, 8, 256, 32, dt / 2, dt)
    drift = 0.0
    for _ in range(1, steps):
        p, _ = orc.step(p, fp, 8, 256, 32, dt, dt)
    p, _ = orc.step(p, fp, 8, 256, 32, dt / 2, 0.0)
    e1 = orc.kinetic_energy(p) + orc.potential_energy(p)
    drift = abs((e1 - e0) / e0)
    assert drift <= 0.05


def test_rejects_bad_input(orc):
    p = ic.random_box(100, 1, 1.0)
    p["pos"][5, 1] = np.nan
    with pytest.raises(ValueError):
        orc.build(p, 8, 64)


def test_baseline_sample_accelerations_match_grouped_pass(orc):
    """The CPU baseline's sampled groups get exactly the grouped pass's
    accelerations (same tree, same walker)."""
    from bhgpu import ic, layout

    p = ic.plummer(20000, 4)
    fp = layout.force_params(0.5, 1e-4, 1.0, True)
    (build_ms, walk_ms, frac), acc, mask = orc.baseline_sample(p, fp, 16, 256, 32, 7, 4, with_acc=True)
    full, _, _, _ = orc.build(p, 16, 256).grouped_pass(fp, 32)
    assert 0.05 < frac < 0.3 and abs(mask.mean() - frac) < 1e-12
    assert np.array_equal(acc[mask], full[mask]) and not acc[~mask].any()


def test_per_node_path_equals_grouped_with_leaf_targets(orc):
    """The per-node path (traversal_ttg.h:156-196: every leaf a target, MAC on
    the leaf's particle box, direct sums over unaccepted leaves) has exactly
    the interaction lists of the grouped walk with target_size 0 (targets =
    leaves), so the same forces up to summation order (eps2 > 0: the self
    term is zero either way)."""
    from bhgpu import ic, layout

    p = ic.plummer(6000, 8)
    for theta, quad in ((0.5, True), (0.7, False), (0.0, True)):
        fp = layout.force_params(theta, 1e-4, 1.0, quad)
        t = orc.build(p, 8, 64)
        a, pp, pc = t.per_node_pass(fp)
        g, gpp, gpc, _ = t.grouped_pass(fp, 0)
        if quad:
            assert np.array_equal(pp, gpp) and np.array_equal(pc, gpc)
        else:  # without quadrupoles accepted cells are monopoles in the P2P list
            assert np.array_equal(pp + pc, gpp)
        assert np.sqrt(((a - g) ** 2).sum() / (g ** 2).sum()) < 1e-13


def _cells_of(keys, depth):
    """Translation (x, y, z) at `depth` of each 63-bit Morton key."""
    def compact(k):
        out = np.zeros_like(k)
        for b in range(21):
            out |= ((k >> np.uint64(3 * b)) & np.uint64(1)) << np.uint64(b)
        return out
    sh = (21 - depth).astype(np.uint64)
    return np.stack([compact(keys >> np.uint64(a)) >> sh for a in range(3)], axis=1).astype(np.int64)


PER_NODE_INPUTS = [("random_box", 3000, 8, 11), ("plummer", 20000, 8, 12), ("hernquist", 12000, 16, 13),
                   ("plummer", 5000, 1, 14)]


@pytest.mark.parametrize("kind,n,ml,seed", PER_NODE_INPUTS)
def test_reference_tree_leaves_equal_flat_leaves(orc, ref, kind, n, ml, seed):
    """The reference's per-node octree (tree.cc, built by the reference's own
    code) has exactly the flat tree's leaves: same level, same cell, same
    particles (no cell reaches the flat tree's depth cap on these inputs)."""
    from bhgpu import ic

    p = ic.random_box(n, seed, 10.0) if kind == "random_box" else getattr(ic, kind)(n, seed)
    level, txyz, count, ids = ref.per_node_tree(p, ml).leaves()
    e = orc.build(p, ml, 64).export()
    nodes = e["nodes"]
    lf = np.flatnonzero(nodes["leaf"] == 1)
    lf = lf[np.argsort(nodes["particle_begin"][lf])]  # Morton order
    assert len(lf) == len(level)
    assert nodes["depth"][lf].max() < 21
    assert np.array_equal(nodes["depth"][lf].astype(np.uint32), level)
    assert np.array_equal(nodes["particle_count"][lf], count)
    first = e["keys"][nodes["particle_begin"][lf]]
    assert np.array_equal(_cells_of(first, nodes["depth"][lf].astype(np.int64)), txyz.astype(np.int64))
    flat_ids = np.concatenate([np.sort(e["particles"]["id"][b:b + c])
                               for b, c in zip(nodes["particle_begin"][lf], nodes["particle_count"][lf])])
    assert np.array_equal(flat_ids, ids)


@pytest.mark.parametrize("kind,n,ml,seed", PER_NODE_INPUTS)
@pytest.mark.parametrize("theta,quad", [(0.5, True), (0.7, False), (0.3, True)])
def test_per_node_pass_matches_reference_tree(orc, ref, kind, n, ml, seed, theta, quad):
    """The oracle's per-node pass (on the flat tree) has, particle by
    particle, the interaction lists of the dispatch recursion
    (traversal_ttg.h:156-196) over the reference's own per-node tree and
    multipoles (tree.cc, upward_ttg.h's moments), and its forces to 1e-12."""
    from bhgpu import ic, layout

    p = ic.random_box(n, seed, 10.0) if kind == "random_box" else getattr(ic, kind)(n, seed)
    fp = layout.force_params(theta, 1e-4, 1.0, quad)
    racc, rpp, rpc = ref.per_node_tree(p, ml).per_node_pass(fp)
    t = orc.build(p, ml, 64)
    oacc, opp, opc = t.per_node_pass(fp)
    ids = t.export()["particles"]["id"]
    assert np.array_equal(rpp[ids], opp) and np.array_equal(rpc[ids], opc)
    err = np.sqrt(((racc[ids] - oacc) ** 2).sum() / (oacc ** 2).sum())
    assert err < 1e-12, err


@pytest.mark.parametrize("name", ["plummer131k"])
def test_oracle_matches_reference_digests(orc, name):
    """At 131,072 particles (several backbone levels above the 4096-particle
    subtrees) the oracle's keys, permutation, node arrays, centres, second
    moments, group order and per-particle list lengths hash to the digests of
    the reference-compiled outputs."""
    import golden_io
    from bhgpu import layout

    d = golden_io.load_digest(name)
    p = golden_io.digest_input(d)
    t = orc.build(p, d["max_leaf"], d["group_size"])
    e = t.export()
    _, npp, npc, _ = t.grouped_pass(layout.force_params(d["theta"], d["eps2"], 1.0, True), d["target_size"])
    got = {k: golden_io.sha(e[k]) for k in ("keys", "perm", "nodes", "center", "iab", "groups")}
    got |= {"n_pp": golden_io.sha(npp), "n_pc": golden_io.sha(npc)}
    assert got == d["sha256"]
