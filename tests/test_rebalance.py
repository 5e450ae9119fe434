"""km1 rebalancer (rebalancer.cpp:36-262), host code shared by the device
pipeline and the oracle: hand-worked cases and, on random overloaded
partitions, balance afterwards, the reported km1 change against a recount,
and moves only out of blocks that were over their bound."""
import numpy as np

import mtk_amd
import oracle as orc


def path4():
    # nets {0,1}, {1,2}, {2,3}, unit weights
    noff = np.array([0, 2, 4, 6], np.uint64)
    pins = np.array([0, 1, 1, 2, 2, 3], np.uint32)
    return orc.Hypergraph(4, noff, pins)


def test_border_vertex_moves_when_km1_does_not_grow():
    # block 0 holds 3 > 2: vertex 2 is its only border vertex; moving it cuts
    # {1,2} and heals {2,3}, gain 0, so it moves at once
    part, delta, moves, bal = path4().rebalance(2, [0, 0, 0, 1], [2, 3])
    assert list(part) == [0, 0, 1, 1] and (delta, moves, bal) == (0, 1, True)


def test_empty_block_repair_then_rebalance():
    # block 1 is empty: vertices 0 and 3 lose the least (one internal net);
    # the tie goes to the larger id, so 3 fills block 1 (+1: {2,3} is cut);
    # block 0 still holds 3 > 2 and vertex 2 moves with gain 0
    part, delta, moves, bal = path4().rebalance(2, [0, 0, 0, 0], [2, 3])
    assert list(part) == [0, 0, 1, 1] and (delta, moves, bal) == (1, 2, True)


def test_balanced_partition_is_left_alone():
    part, delta, moves, bal = path4().rebalance(2, [0, 0, 1, 1], [2, 2])
    assert list(part) == [0, 0, 1, 1] and (delta, moves, bal) == (0, 0, True)


def test_random_overloaded_partitions():
    for seed in range(6):
        n, k = 3000, 4 + seed % 3
        h = mtk_amd.HostHypergraph.powerlaw(n, n, 2.3, 2, 200, window=512, local_fraction=0.8,
                                            vertex_weight_min=1, vertex_weight_max=3, seed=100 + seed)
        a = h.arrays()
        h.close()
        oh = orc.Hypergraph(n, a["noff"], a["pins"], a["vw"], a["nw"])
        rng = np.random.default_rng(seed)
        part = rng.integers(0, k, n).astype(np.int32)
        part[rng.random(n) < 0.5] = 0  # block 0 far over its bound
        W = int(a["vw"].sum())
        maxw = int(1.03 * -(-W // k))
        before = oh.phg_init(k, part)
        over = np.bincount(part, weights=a["vw"], minlength=k) > maxw
        new, delta, moves, bal = oh.rebalance(k, part, [maxw] * k)
        after = oh.phg_init(k, new)
        assert bal and moves > 0
        assert np.bincount(new, weights=a["vw"], minlength=k).max() <= maxw
        assert after["km1"] - before["km1"] == delta
        moved = new != part
        assert over[part[moved]].all()  # only vertices of overloaded blocks move
