"""Steps on either side of the multilevel partitioner in partition()
(partitioner.cpp:37-38, 74-86, 190-195): degree-zero vertex removal and its
bin-packing restore (degree_zero_hn_remover.h:46-83) and large-net removal
(large_he_remover.h:47-92).  The oracle is pinned here by hand-worked
examples; the device path is compared with it in test_gpu_sanitize.py."""
import numpy as np

import mtk_amd
import oracle as orc

REMOVED = 0xFFFFFFFF


def small():
    # vertices 4 and 7 have no nets; net 1 has five pins
    noff = np.array([0, 2, 7, 9], np.uint64)
    pins = np.array([0, 1, 1, 2, 3, 5, 6, 3, 5], np.uint32)
    vw = np.arange(1, 9, dtype=np.int32)
    nw = np.array([10, 20, 30], np.int32)
    return orc.Hypergraph(8, noff, pins, vw, nw)


def test_large_net_threshold():
    # max(max(N * 0.01, 100), 50000), truncated to an integer
    for n, want in [(0, 50000), (1000, 50000), (4_000_000, 50000), (20_000_000, 200000),
                    (123_456_789, 1234567)]:
        assert orc.large_net_threshold(n) == want
        assert mtk_amd.large_net_threshold(n) == want


def test_reduce_removes_first_degree_zero_vertices_and_large_nets():
    h = small()
    # contraction limit 7: one vertex may go, the first without nets (4)
    r, vmap, rv, rn = h.reduce(7, 4)
    a = r.arrays()
    assert (rv, rn) == (1, 1)
    assert list(vmap) == [0, 1, 2, 3, REMOVED, 4, 5, 6]
    assert list(a["noff"]) == [0, 2, 4]
    assert list(a["pins"]) == [0, 1, 3, 4]
    assert list(a["vw"]) == [1, 2, 3, 4, 6, 7, 8]
    assert list(a["nw"]) == [10, 30]
    # contraction limit 5: both vertices without nets go (vertex 6 keeps its
    # degree of one: degrees are those of the input)
    r, vmap, rv, rn = h.reduce(5, 4)
    assert (rv, rn) == (2, 1)
    assert list(vmap) == [0, 1, 2, 3, REMOVED, 4, 5, REMOVED]
    assert list(r.arrays()["pins"]) == [0, 1, 3, 4]
    # at the contraction limit nothing is removed but the large net
    r, vmap, rv, rn = h.reduce(8, 4)
    assert (rv, rn) == (0, 1)
    assert list(vmap) == list(range(8))
    # threshold at the net size: the net stays (removal needs |e| > threshold)
    r, vmap, rv, rn = h.reduce(8, 5)
    assert (rv, rn) == (0, 0)


def test_bin_packing_hand_worked():
    # removed 0..3 of weights 5, 3, 3, 1 into blocks of weights 10, 4, 7:
    # 0 -> block 1 (9, list 2 1 0), 1 -> block 2 (10, list 1 2 0),
    # 2 -> block 1 (12, list 2 0 1), 3 -> block 2 (11, list 0 2 1)
    part, pw = orc.restore_degree_zero([0, 1, 2, 3], [5, 3, 3, 1], 3, [10, 4, 7], [-1] * 4)
    assert list(part) == [1, 2, 1, 2]
    assert list(pw) == [10, 12, 11]
    # equal block weights: the smaller block id first; a block that is not
    # heavier than the next stays in front
    part, pw = orc.restore_degree_zero([0, 1], [2, 2], 2, [3, 3], [-1, -1])
    assert list(part) == [0, 1] and list(pw) == [5, 5]
    # heavier vertices first, equal weights in id order (the removed list is
    # ascending): 2 -> block 0 (4, list 1 0), 1 -> block 1 (1), 3 -> block 1
    # (2); vertex 0 was not removed and keeps its block
    part, pw = orc.restore_degree_zero([1, 2, 3], [0, 1, 4, 1], 2, [0, 0], [7, -1, -1, -1])
    assert list(part) == [7, 1, 0, 1] and list(pw) == [4, 2]


def test_pipeline_assigns_removed_vertices_within_balance():
    # a power-law hypergraph with 40% extra vertices that have no nets
    n0, extra, k, eps = 3000, 2000, 4, 0.03
    h = mtk_amd.HostHypergraph.powerlaw(n0, n0, 2.3, 2, 200, window=512, local_fraction=0.8, seed=11)
    a = h.arrays()
    h.close()
    n = n0 + extra
    rng = np.random.default_rng(3)
    vw = np.concatenate([a["vw"], rng.integers(1, 4, extra).astype(np.int32)])
    deg = np.bincount(a["pins"].astype(np.int64), minlength=n)
    oh = orc.Hypergraph(n, a["noff"], a["pins"], vw, a["nw"])
    part, km1, st = oh.partition(k, eps, 5, True)
    zero = int((deg == 0).sum())
    assert st["removed_vertices"] == min(zero, n - 160 * k) and st["removed_vertices"] > 0
    assert st["removed_nets"] == 0
    assert part.min() >= 0 and part.max() < k
    perfect = -(-int(vw.sum()) // k)
    assert np.bincount(part, weights=vw, minlength=k).max() <= int((1.0 + eps) * perfect)
