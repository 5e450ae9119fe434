"""The oracle's initial partitioner (oracle/initial.cpp), the restatement the
device initial partitioner is checked against.  Pinned by hand-worked growing
orders (greedy scores with their tie rule, level-synchronous BFS in ascending
id order), the Feistel visiting order being a permutation, incremental gain
and pin-count upkeep against recounts after every move, brute-force optimal
bisections of tiny hypergraphs, and balance / km1 / determinism properties of
the whole recursion."""
import itertools

import numpy as np
import pytest

import mtk_amd
import oracle as orc
from helpers import csr


def hg(n, nets, vw=None, nw=None):
    noff, pins = csr(nets)
    vw = np.ones(n, np.int32) if vw is None else np.asarray(vw, np.int32)
    nw = np.ones(len(nets), np.int32) if nw is None else np.asarray(nw, np.int32)
    return orc.Hypergraph(n, noff, pins, vw, nw)


def stream_with_start(r, n, start):
    """A stream key whose run r starts growing at `start`."""
    for s in range(100000):
        if orc.ip_start(orc.ip_run_key(s, r), n) == start:
            return s
    raise AssertionError("no stream found")


def recount_cut(nets, side, nw=None):
    c = 0
    for i, e in enumerate(nets):
        if len({int(side[v]) for v in e}) == 2:
            c += 1 if nw is None else int(nw[i])
    return c


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 9, 16, 17, 100, 255, 256, 257, 1000, 4097])
def test_visiting_order_is_a_permutation(n):
    for key in (0, 1, 0xDEADBEEF, 2**63 + 5):
        p = orc.ip_perm(key, n)
        assert sorted(p.tolist()) == list(range(n))
    if n >= 16:
        a, b = orc.ip_perm(1, n), orc.ip_perm(2, n)
        assert not np.array_equal(a, b)
        assert not np.array_equal(a, np.arange(n))


def test_greedy_growing_order_hand_worked():
    # from vertex 0: e0 gives 1 the score 2^16, e1 gives 2 and 3 2^15 each;
    # 1 joins and e2 gives 4 2^16; 4 joins (2^16 > 2^15) and e3 gives 5 2^16;
    # 5 joins and side 0 reaches its target of 4
    nets = [[0, 1], [0, 2, 3], [1, 4], [4, 5], [2, 3]]
    h = hg(6, nets)
    s = stream_with_start(0, 6, 0)
    side, cut, bal, w0 = h.ip_bisect_run(4, 4, 4, s, 0, passes=0)
    assert set(np.flatnonzero(side == 0).tolist()) == {0, 1, 4, 5}
    assert w0 == 4 and bal and cut == 1 == recount_cut(nets, side)


def test_greedy_growing_ties_go_to_the_smaller_id():
    nets = [[0, 2], [0, 1], [3, 4]]
    h = hg(5, nets)
    s = stream_with_start(0, 5, 0)
    side, _, _, _ = h.ip_bisect_run(2, 5, 2, s, 0, passes=0)
    assert set(np.flatnonzero(side == 0).tolist()) == {0, 1}


def test_greedy_growing_ignores_nets_over_100_pins():
    # the 101-pin net does not score; 1 is reached through the 2-pin net only
    big = list(range(0, 101))
    nets = [big, [0, 150]]
    h = hg(200, nets)
    s = stream_with_start(0, 200, 0)
    side, _, _, _ = h.ip_bisect_run(2, 200, 2, s, 0, passes=0)
    assert set(np.flatnonzero(side == 0).tolist()) == {0, 150}


def test_bfs_levels_join_in_ascending_id_order():
    # level 0 {3}, level 1 {4, 5}, level 2 {0, 6} (6 is found first, through
    # 4, but 0 has the smaller id); a target of 4 stops after 0
    nets = [[3, 4], [3, 5], [4, 6], [5, 0], [1, 2]]
    h = hg(7, nets)
    s = stream_with_start(1, 7, 3)
    side, _, _, w0 = h.ip_bisect_run(4, 7, 4, s, 1, passes=0)
    assert set(np.flatnonzero(side == 0).tolist()) == {0, 3, 4, 5} and w0 == 4
    side, _, _, _ = h.ip_bisect_run(5, 7, 5, s, 1, passes=0)
    assert set(np.flatnonzero(side == 0).tolist()) == {0, 3, 4, 5, 6}


def test_bfs_new_seed_after_the_component_is_exhausted():
    # component {3, 4} is exhausted; the next unseen vertex after the start
    # (cyclically: 5, then 6, 0, ...) seeds the next level
    nets = [[3, 4], [5, 6], [0, 1]]
    h = hg(7, nets)
    s = stream_with_start(1, 7, 3)
    side, _, _, _ = h.ip_bisect_run(3, 7, 3, s, 1, passes=0)
    assert set(np.flatnonzero(side == 0).tolist()) == {3, 4, 5}


def mix64(z):
    """SplitMix64 finalizer (Steele, Lea, Flood, OOPSLA 2014)."""
    M = (1 << 64) - 1
    z = (z + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_random_fill_walks_the_permutation_skipping_heavy_vertices():
    n = 40
    rng = np.random.default_rng(3)
    vw = rng.integers(1, 9, n).astype(np.int32)
    h = hg(n, [[i, (i + 1) % n] for i in range(n)], vw=vw)
    W = int(vw.sum())
    target, mx = W // 2, W // 2 + 3
    s = 12345
    side, _, _, w0 = h.ip_bisect_run(mx, W, target, s, 2, passes=0)
    order = orc.ip_perm(mix64(orc.ip_run_key(s, 2) ^ 0x13198A2E03707344), n)
    want, acc = set(), 0
    for u in order:
        if acc >= target:
            break
        if acc + vw[u] > mx:
            continue
        want.add(int(u))
        acc += int(vw[u])
    assert set(np.flatnonzero(side == 0).tolist()) == want and w0 == acc


@pytest.mark.parametrize("seed", range(6))
def test_incremental_gains_and_pin_counts_equal_recounts(seed):
    rng = np.random.default_rng(seed)
    n, m = 60, 90
    nets = []
    for _ in range(m):
        s = int(rng.integers(2, 12))
        nets.append(rng.choice(n, size=s, replace=False).tolist())
    nw = rng.integers(1, 5, m).astype(np.int32)
    h = hg(n, nets, nw=nw)
    side = rng.integers(0, 2, n).astype(np.uint8)
    moves = rng.integers(0, n, 200).astype(np.uint32)
    assert h.ip_move_check(side, moves) == 0


def brute_force_best_cut(n, nets, vw, max0, max1):
    best = None
    for bits in itertools.product((0, 1), repeat=n):
        side = np.array(bits, np.uint8)
        w0 = int(vw[side == 0].sum())
        if w0 > max0 or int(vw.sum()) - w0 > max1:
            continue
        c = recount_cut(nets, side)
        best = c if best is None else min(best, c)
    return best


@pytest.mark.parametrize("seed", range(8))
def test_runs_are_never_better_than_the_optimum_and_usually_reach_it(seed):
    rng = np.random.default_rng(100 + seed)
    n = 11
    nets = [rng.choice(n, size=int(rng.integers(2, 5)), replace=False).tolist() for _ in range(16)]
    vw = np.ones(n, np.int32)
    h = hg(n, nets)
    max0, max1, target0 = 6, 6, 6
    opt = brute_force_best_cut(n, nets, vw, max0, max1)
    cuts = []
    for r in range(20):
        side, cut, bal, w0 = h.ip_bisect_run(max0, max1, target0, 99 + seed, r)
        assert cut == recount_cut(nets, side)
        assert w0 == int(vw[side == 0].sum())
        assert bal == (w0 <= max0 and n - w0 <= max1)
        if bal:
            assert cut >= opt
            cuts.append(cut)
    assert min(cuts) == opt


def test_two_cliques_are_separated():
    a, b = list(range(8)), list(range(8, 16))
    nets = [list(p) for p in itertools.combinations(a, 2)] + [list(p) for p in itertools.combinations(b, 2)]
    nets.append([7, 8])
    h = hg(16, nets)
    part, km1 = h.initial_partition(2, [8, 8], 42)
    assert km1 == 1
    assert len(set(part[:8])) == 1 and len(set(part[8:])) == 1


@pytest.mark.parametrize("k", [2, 3, 5, 8, 16])
def test_recursion_balanced_deterministic_and_thread_independent(k):
    h = mtk_amd.HostHypergraph.powerlaw(3000, 3600, 2.4, 2, 300, window=256, local_fraction=0.8,
                                        vertex_weight_min=1, vertex_weight_max=6, seed=11 + k)
    a = h.arrays()
    h.close()
    oh = orc.Hypergraph(a["n"], a["noff"], a["pins"], a["vw"], a["nw"])
    W = int(a["vw"].sum())
    maxw = int(1.03 * -(-W // k))
    p1, km1 = oh.initial_partition(k, [maxw] * k, 7, threads=1)
    p4, km4 = oh.initial_partition(k, [maxw] * k, 7, threads=4)
    assert np.array_equal(p1, p4) and km1 == km4
    assert p1.min() >= 0 and p1.max() < k and len(set(p1.tolist())) == k
    bw = np.bincount(p1, weights=a["vw"], minlength=k)
    assert bw.max() <= maxw
    ref = oh.phg_init(k, p1)
    assert ref["km1"] == km1
    p2, _ = oh.initial_partition(k, [maxw] * k, 8, threads=4)
    assert not np.array_equal(p1, p2)


def test_greedy_growing_takes_n_over_2048_vertices_per_step():
    # 4096 vertices: two picks per step.  From 0, the scores are 1: 3 * 2^16,
    # 2: 2 * 2^16, 3: 2^16; one pick per step would take 1 and then 5 (its
    # net of weight 100 makes 5 the best), two picks take 1 and 2 together
    n = 4096
    nets = [[0, 1], [0, 2], [0, 3], [1, 5]]
    h = hg(n, nets, nw=[3, 2, 1, 100])
    s = stream_with_start(0, n, 0)
    side, _, _, w0 = h.ip_bisect_run(3, n, 3, s, 0, passes=0)
    assert set(np.flatnonzero(side == 0).tolist()) == {0, 1, 2} and w0 == 3
    h1 = hg(2000, nets, nw=[3, 2, 1, 100])
    s1 = stream_with_start(0, 2000, 0)
    side, _, _, _ = h1.ip_bisect_run(3, 2000, 3, s1, 0, passes=0)
    assert set(np.flatnonzero(side == 0).tolist()) == {0, 1, 5}


def test_no_random_fill_above_1024_vertices():
    # run 2 fills at random up to 1024 vertices and grows greedily beyond
    for n, random in ((1024, True), (1025, False)):
        vw = np.ones(n, np.int32)
        h = hg(n, [[i, i + 1] for i in range(n - 1)], vw=vw)
        target = n // 2
        s = 4242
        side, _, _, _ = h.ip_bisect_run(target, n, target, s, 2, passes=0)
        order = orc.ip_perm(mix64(orc.ip_run_key(s, 2) ^ 0x13198A2E03707344), n)
        fill = set(order[:target].tolist())
        assert (set(np.flatnonzero(side == 0).tolist()) == fill) == random
