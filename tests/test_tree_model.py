"""The GPU tree construction's algorithm (tests/tree_model.py mirrors the
kernels of bh_tree.hip) against the oracle's reference-layout tree, on the
CPU: same nodes, children, counts, depths, leaves and group order after the
reference renumbering, for every layout regime and degenerate input."""
import numpy as np
import pytest

import tree_model as tm
from bhgpu import ic


def compare(orc, p, ml, gs):
    oe = orc.build(p, ml, gs).export()
    t = tm.build_bfs(oe["keys"], ml)
    ref_of, groups = tm.reference_order(t, gs)
    on = oe["nodes"]
    assert len(t["begin"]) == len(on)
    assert sorted(ref_of.tolist()) == list(range(len(on)))
    assert np.array_equal(on["particle_begin"][ref_of], t["begin"])
    assert np.array_equal(on["particle_count"][ref_of], t["count"])
    assert np.array_equal(on["depth"][ref_of], t["depth"])
    assert np.array_equal(on["leaf"][ref_of].astype(bool), t["leaf"])
    assert np.array_equal(on["num_children"][ref_of], t["nchild"])
    assert np.array_equal(on["first_child"][ref_of], np.where(t["leaf"], 0, ref_of[t["first_child"]]))
    assert np.array_equal(oe["groups"], groups)


@pytest.mark.parametrize("n,ml,gs,seed", [(1, 8, 64, 1), (5, 8, 64, 1), (9, 8, 64, 2), (200, 4, 16, 1),
                                          (2000, 8, 64, 4), (20000, 16, 64, 5), (20000, 16, 8192, 5),
                                          (5000, 1, 64, 6), (3000, 0, 64, 7), (4096, 16, 256, 8), (4097, 16, 256, 8)])
def test_model_matches_oracle(orc, n, ml, gs, seed):
    compare(orc, ic.random_box(n, seed, 10.0), ml, gs)


def test_model_plummer(orc):
    compare(orc, ic.plummer(30000, 3), 32, 256)


def test_model_duplicates(orc):
    p = ic.random_box(300, 3, 10.0)
    p["pos"][:100] = p["pos"][0]
    compare(orc, p, 4, 16)
    q = ic.random_box(50, 3, 10.0)
    q["pos"][:] = q["pos"][0]
    compare(orc, q, 4, 16)


def test_leaf_depth_window_identity():
    # the cell of depth d containing i holds > max_leaf particles iff d <= W_i
    rng = np.random.default_rng(1)
    keys = np.sort(rng.integers(0, 1 << 20, size=500).astype(np.uint64) << np.uint64(40))
    for ml in (0, 1, 3, 8):
        L, H = tm.leaf_depths(keys, ml)
        for i in range(0, 500, 7):
            for d in range(22):
                shift = np.uint64(63 - 3 * d)
                same = (keys >> shift) == (keys[i] >> shift)
                big = same.sum() > ml
                if d <= 21:
                    # H_i is the first depth whose cell is not split
                    assert (d < H[i]) == (big and d < 21)
