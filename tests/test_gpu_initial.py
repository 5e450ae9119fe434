"""Device initial partitioning (csrc/initial.hip) against the oracle's
restatement (oracle/initial.cpp): every run of a bisection (side of every
vertex, cut, balance, side weights) and the whole recursive bisection
(block of every vertex, km1) are bit-exact: growing from precomputed clique
scores (up to 4,096 vertices) and by walking the nets, per-vertex state in
LDS (up to 8,192 vertices), sides in LDS with scores in global memory (up to
32,768) and all in global memory, 16-bit and 32-bit pin counts."""
import math

import numpy as np
import pytest

import mtk_amd
import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = mtk_amd.Context(0)
    yield c
    c.close()


def powerlaw(n, m, max_size, weighted, seed):
    h = mtk_amd.HostHypergraph.powerlaw(n, m, 2.3, 2, max_size, window=512, local_fraction=0.8,
                                        vertex_weight_min=1, vertex_weight_max=6 if weighted else 1, seed=seed)
    a = h.arrays()
    return h, a


CASES = {
    # nets above 64 pins go through the wave-per-net queue of a move, nets
    # above 100 pins do not steer growing
    "small": (500, 800, 150, False, 1),
    "weighted": (3000, 3600, 300, True, 2),
    "mid": (6000, 7000, 200, True, 6),  # growing walks the nets (no clique-score matrix)
    "global": (9500, 9000, 60, True, 3),  # scores and gains in global memory
    "huge": (40000, 36000, 60, True, 8),  # all per-vertex state in global memory
    "dense": (300, 6000, 40, False, 4),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("limits", ["half", "third", "repair"])
def test_bisection_runs_equal_oracle(ctx, case, limits):
    n, m, mx, weighted, seed = CASES[case]
    h, a = powerlaw(n, m, mx, weighted, seed)
    W = int(a["vw"].sum())
    if limits == "half":
        max0 = max1 = int(1.03 * math.ceil(W / 2))
        target0 = math.ceil(W / 2)
    elif limits == "third":
        max0, max1, target0 = int(1.02 * math.ceil(W / 3)), int(1.02 * math.ceil(2 * W / 3)), math.ceil(W / 3)
    else:  # side 1 starts far over its bound and is repaired
        max0, max1, target0 = W, W // 2, W // 4
    stream = orc.ip_stream(seed, 0, 2)
    print("case", case, limits, "n", n, "pins", len(a["pins"]), flush=True)
    dh = mtk_amd.Hypergraph(ctx, host=h)
    sides, res = dh.ip_bisect_runs(max0, max1, target0, stream, runs=12, passes=4)
    oh = orc.Hypergraph(a["n"], a["noff"], a["pins"], a["vw"], a["nw"])
    for r in range(12):
        side, cut, bal, w0 = oh.ip_bisect_run(max0, max1, target0, stream, r, passes=4)
        assert np.array_equal(sides[r], side), "run %d" % r
        assert list(res[r]) == [cut, int(bal), w0, W - w0], "run %d" % r
    dh.close(); h.close()


def test_bisection_runs_with_a_net_of_more_than_65535_pins(ctx):
    # pin counts no longer fit 16 bits: the two-word count layout
    n = 70000
    rng = np.random.default_rng(5)
    nets = [rng.choice(n, size=int(rng.integers(2, 9)), replace=False) for _ in range(30000)]
    nets.append(np.arange(0, n, 1)[:66000])
    noff = np.zeros(len(nets) + 1, np.uint64)
    noff[1:] = np.cumsum([len(e) for e in nets])
    pins = np.concatenate(nets).astype(np.uint32)
    vw = rng.integers(1, 4, n).astype(np.int32)
    nw = np.ones(len(nets), np.int32)
    W = int(vw.sum())
    max0, max1, target0 = W // 40, W, W // 50
    stream = orc.ip_stream(9, 0, 2)
    dh = mtk_amd.Hypergraph(ctx, n=n, noff=noff, pins=pins, vw=vw, nw=nw)
    sides, res = dh.ip_bisect_runs(max0, max1, target0, stream, runs=6, passes=2)
    oh = orc.Hypergraph(n, noff, pins, vw, nw)
    for r in range(6):
        side, cut, bal, w0 = oh.ip_bisect_run(max0, max1, target0, stream, r, passes=2)
        assert np.array_equal(sides[r], side), "run %d" % r
        assert list(res[r]) == [cut, int(bal), w0, W - w0], "run %d" % r
    dh.close()


def test_bisection_single_vertex_and_no_nets(ctx):
    for n, nets in ((1, []), (40, [])):
        noff = np.zeros(1, np.uint64)
        pins = np.zeros(0, np.uint32)
        vw = np.arange(1, n + 1, dtype=np.int32)
        nw = np.zeros(0, np.int32)
        dh = mtk_amd.Hypergraph(ctx, n=n, noff=noff, pins=pins, vw=vw, nw=nw)
        W = int(vw.sum())
        max0 = max1 = int(1.03 * math.ceil(W / 2))
        sides, res = dh.ip_bisect_runs(max0, max1, math.ceil(W / 2), 77, runs=6, passes=4)
        oh = orc.Hypergraph(n, noff, pins, vw, nw)
        for r in range(6):
            side, cut, bal, w0 = oh.ip_bisect_run(max0, max1, math.ceil(W / 2), 77, r, passes=4)
            assert np.array_equal(sides[r], side) and list(res[r]) == [cut, int(bal), w0, W - w0]
        dh.close()


@pytest.mark.parametrize("k,seed", [(2, 11), (3, 12), (4, 13), (7, 14), (16, 15), (64, 16), (128, 17)])
def test_initial_partition_equals_oracle(ctx, k, seed):
    n = 4000
    h, a = powerlaw(n, 5000, 400, True, seed)
    eps = 0.03
    W = int(a["vw"].sum())
    perfect = math.ceil(W / k)
    maxw = int((1.0 + eps) * perfect)
    print("k", k, "seed", seed, flush=True)
    dh = mtk_amd.Hypergraph(ctx, host=h)
    dp = mtk_amd.Partition(dh, k)
    km1 = dp.initial_partition(eps, seed, runs=20)
    oh = orc.Hypergraph(a["n"], a["noff"], a["pins"], a["vw"], a["nw"])
    opart, okm1 = oh.initial_partition(k, [maxw] * k, seed, runs=20, threads=8)
    assert np.array_equal(dp.parts(), opart)
    assert km1 == okm1 == dp.km1()
    bw = np.bincount(opart, weights=a["vw"], minlength=k)
    assert bw.max() <= maxw
    dp.close(); dh.close(); h.close()
