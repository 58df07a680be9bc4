"""Routing schedule: inflow/outflow cells and rout_order (rout_prepare.cpp:338-439, 619-670,
834-886) -- bit-exact against the oracle and against hand-derived known answers."""
import numpy as np
import pytest


def graph(n, edges):
    """inflow9/outflow arrays for a list of (upstream, downstream) 1-based edges."""
    inflow9 = np.zeros((n, 9), dtype=np.int32)
    outflow = np.zeros(n, dtype=np.int32)
    slot = np.zeros(n, dtype=int)
    for u, d in edges:
        outflow[u - 1] = d
        inflow9[d - 1, slot[d - 1]] = u
        slot[d - 1] += 1
    return inflow9, outflow


KNOWN = [
    # chain 1 -> 2 -> 3 -> 4: one cell per sweep
    (4, [(1, 2), (2, 3), (3, 4)], [1, 2, 3, 4], [0, 1, 2, 3]),
    # chain listed downstream-first: 4 -> 3 -> 2 -> 1
    (4, [(4, 3), (3, 2), (2, 1)], [4, 3, 2, 1], [3, 2, 1, 0]),
    # star: 1,2,3,5 all drain to 4; sweep 0 numbers 1,2,3,5 in cell order
    (5, [(1, 4), (2, 4), (3, 4), (5, 4)], [1, 2, 3, 5, 4], [0, 0, 0, 1, 0]),
    # Y-tree with unequal arms: 1->2->5, 3->5, 5->4, 6 isolated
    (6, [(1, 2), (2, 5), (3, 5), (5, 4)], [1, 4, 2, 6, 5, 3], [0, 1, 0, 3, 2, 0]),
]


@pytest.mark.parametrize("n,edges,order,level", KNOWN)
def test_known_answers(wg, oracle, n, edges, order, level):
    inflow9, outflow = graph(n, edges)
    o, l = wg.rout_order(inflow9, outflow)
    assert o.tolist() == order and l.tolist() == level
    rc, o2, l2 = oracle.rout_order(inflow9, outflow)
    assert rc == 0 and o2.tolist() == order and l2.tolist() == level


def test_cycle_is_rejected(wg, oracle):
    inflow9, outflow = graph(4, [(1, 2), (2, 3), (3, 2), (4, 1)])
    with pytest.raises(wg.WaterGapError) as e:
        wg.rout_order(inflow9, outflow)
    assert e.value.code == -2
    rc, _, _ = oracle.rout_order(inflow9, outflow)
    assert rc == -2


def test_inconsistent_outflow(wg, oracle):
    # Cell 2 lists only cell 3 as inflow, but cells 1 and 3 both drain to it: in
    # the first sweep its counter goes 1 -> 0 -> -1, never returns to 0, and the
    # cell is never numbered (rout_prepare.cpp:876-881).
    inflow9 = np.zeros((3, 9), dtype=np.int32)
    inflow9[1, 0] = 3
    outflow = np.array([2, 0, 2], dtype=np.int32)
    with pytest.raises(wg.WaterGapError):
        wg.rout_order(inflow9, outflow)
    assert oracle.rout_order(inflow9, outflow)[0] == -2
    # An outflow cell that simply does not list its inflow is numbered in the same
    # sweep and is not an error in the reference.
    inflow9 = np.zeros((3, 9), dtype=np.int32)
    inflow9[2, 0] = 2
    outflow = np.array([2, 3, 0], dtype=np.int32)
    o, l = wg.rout_order(inflow9, outflow)
    rc, o2, l2 = oracle.rout_order(inflow9, outflow)
    assert rc == 0 and o.tolist() == o2.tolist() == [1, 2, 3] and l.tolist() == l2.tolist()


def test_brute_force_random_forests(wg, oracle):
    rng = np.random.default_rng(7)
    for trial in range(50):
        n = int(rng.integers(1, 60))
        # random forest: each cell drains to a higher-numbered cell or is an outlet
        perm = rng.permutation(n) + 1
        edges = []
        for k in range(n - 1):
            if rng.random() < 0.8:
                edges.append((int(perm[k]), int(perm[rng.integers(k + 1, n)])))
        # keep at most 8 inflows per cell
        cnt = {}
        kept = []
        for u, d in edges:
            if cnt.get(d, 0) < 8:
                kept.append((u, d))
                cnt[d] = cnt.get(d, 0) + 1
        inflow9, outflow = graph(n, kept)
        o, l = wg.rout_order(inflow9, outflow)
        rc, o2, l2 = oracle.rout_order(inflow9, outflow)
        assert rc == 0
        assert np.array_equal(o, o2) and np.array_equal(l, l2)
        # brute force: level = longest upstream path; order sorted by (level, index)
        lev = np.zeros(n, dtype=int)
        for _ in range(n):
            for u, d in kept:
                lev[d - 1] = max(lev[d - 1], lev[u - 1] + 1)
        assert np.array_equal(l, lev)
        expect = np.lexsort((np.arange(n), lev))
        assert np.array_equal(np.argsort(o), expect)


@pytest.mark.parametrize("which", ["regional_world", "regional_edge_world", "global_world"])
def test_worlds_match_oracle(wg, oracle, request, which):
    w = request.getfixturevalue(which)
    i9, of = wg.inflow_outflow(w)
    i9o, ofo = oracle.inflow_outflow(w)
    assert np.array_equal(i9, i9o) and np.array_equal(of, ofo)
    assert np.array_equal(of, w.downstream)
    rc, order, level = oracle.rout_order(i9o, ofo)
    assert rc == 0
    assert np.array_equal(order, w.rout_order) and np.array_equal(level, w.level)
