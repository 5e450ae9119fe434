"""The oracle's FM pieces pinned by what the mathematics fixes:
recalculated gains against brute-force km1 differences of the sequence
applied move by move (global_rollback.cpp:207-283 must be exact), the best
prefix against an enumeration of all prefixes, and the refinement against
its own invariants (the reported improvement is the km1 decrease, the
blocks stay within their bounds, a second run from a local optimum of one
round finds nothing)."""
import numpy as np
import pytest

import mtk_amd
import oracle as orc


def instance(n, m, max_size, seed, k, weighted=False):
    h = mtk_amd.HostHypergraph.powerlaw(n, m, 2.5, 2, max_size, window=max(16, n // 8), local_fraction=0.8,
                                        seed=seed)
    a = h.arrays()
    h.close()
    rng = np.random.default_rng(seed)
    vw = rng.integers(1, 6, n).astype(np.int32) if weighted else None
    nw = rng.integers(1, 5, len(a["noff"]) - 1).astype(np.int32) if weighted else None
    return orc.Hypergraph(n, a["noff"], a["pins"], vw, nw), vw, rng


@pytest.mark.parametrize("k,seed,weighted", [(2, 1, False), (4, 2, True), (9, 3, True), (70, 4, False)])
def test_recalculated_gains_are_km1_differences(k, seed, weighted):
    n = 400
    oh, vw, rng = instance(n, 500, 30, seed, k, weighted)
    part = rng.integers(0, k, n).astype(np.int32)
    for _ in range(20):
        nm = int(rng.integers(1, 60))
        v = rng.choice(n, nm, replace=False).astype(np.uint32)
        frm = part[v]
        to = ((frm + rng.integers(1, k, nm)) % k).astype(np.int32)
        g = oh.fm_recalculate(k, part, v, frm, to)
        cur = part.copy()
        prev = oh.phg_init(k, cur)["km1"]
        for i in range(nm):
            cur[v[i]] = to[i]
            now = oh.phg_init(k, cur)["km1"]
            assert prev - now == g[i], (i, prev - now, g[i])
            prev = now


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_best_prefix_by_enumeration(seed):
    n, k = 200, 3
    oh, vw, rng = instance(n, 250, 20, seed, k, True)
    w = vw.astype(np.int64)
    part = rng.integers(0, k, n).astype(np.int32)
    bw0 = np.bincount(part, weights=w, minlength=k).astype(np.int64)
    for _ in range(30):
        nm = int(rng.integers(1, 40))
        v = rng.choice(n, nm, replace=False).astype(np.uint32)
        frm = part[v]
        to = ((frm + rng.integers(1, k, nm)) % k).astype(np.int32)
        gain = rng.integers(-3, 4, nm).astype(np.int64)
        max_w = (bw0.max() + rng.integers(-20, 20, k)).astype(np.int64)
        length, best = oh.fm_best_prefix(k, part, v, frm, to, gain, max_w)
        # every prefix with no block above its bound, the empty one included;
        # best gain, then the lighter heaviest block, then the shorter prefix
        cands = [(0, int(bw0.max()), 0)]
        bw = bw0.copy()
        s = 0
        for i in range(nm):
            s += int(gain[i])
            bw[frm[i]] -= w[v[i]]
            bw[to[i]] += w[v[i]]
            if np.all(bw <= max_w):
                cands.append((s, int(bw.max()), i + 1))
        cands.sort(key=lambda c: (-c[0], c[1], c[2]))
        assert (length, best) == (cands[0][2], cands[0][0])


@pytest.mark.parametrize("k,weighted", [(2, False), (5, True), (16, True)])
def test_refinement_invariants(k, weighted):
    n = 3000
    oh, vw, rng = instance(n, 3500, 80, 5, k, weighted)
    w = np.ones(n, np.int64) if vw is None else vw.astype(np.int64)
    perfect = -(-int(w.sum()) // k)
    maxw = [int(1.03 * perfect)] * k
    # contiguous blocks by cumulative weight: every block within its bound
    part = np.minimum((np.cumsum(w) - w) * k // int(w.sum()), k - 1).astype(np.int32)
    assert np.bincount(part, weights=w, minlength=k).max() <= maxw[0]
    start = oh.phg_init(k, part)["km1"]
    p1, km1, st = oh.fm_refine(k, part, maxw, seed=3)
    assert start - km1 == st["improvement"] > 0
    assert np.bincount(p1, weights=w, minlength=k).max() <= maxw[0]
    assert st["rounds"] >= 1 and st["moves"] > 0
    # the same partition again: a round that improves nothing applies nothing
    p2, km2, st2 = oh.fm_refine(k, p1, maxw, seed=3, rounds=1)
    assert km1 - km2 == st2["improvement"] >= 0
    if st2["improvement"] == 0:
        assert np.array_equal(p1, p2)


def test_no_border_vertices():
    n, k = 100, 2
    oh, vw, rng = instance(n, 120, 10, 6, k)
    part = np.zeros(n, np.int32)  # every net inside block 0: no border vertex
    p1, km1, st = oh.fm_refine(k, part, [n, n], seed=1)
    assert km1 == 0 and st["rounds"] == 0 and np.array_equal(p1, part)
