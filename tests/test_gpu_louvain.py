"""Parity of the HIP Louvain (star expansion, all levels) with the oracle's
synchronous restatement.  With integer-valued arc weights every FP64 sum is
exact, so community ids and round statistics are compared bit for bit."""
import os

import numpy as np
import pytest

import mtk_amd
import oracle as orc
from helpers import GOLDEN, csr, read_hgr, vectors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = mtk_amd.Context(0)
    yield c
    c.close()


def powerlaw(n, m, max_size, seed, window=512, exponent=2.3, heavy_nets=0, huge_nets=0):
    h = mtk_amd.HostHypergraph.powerlaw(n, m, exponent, 2, max_size, window=window, local_fraction=0.8, seed=seed)
    a = h.arrays()
    h.close()
    nets = [a["pins"][a["noff"][e]:a["noff"][e + 1]].tolist() for e in range(len(a["noff"]) - 1)]
    rng = np.random.default_rng(seed)
    for _ in range(heavy_nets):  # net nodes of degree > 1024: a 1024-thread block each
        nets.append(sorted(rng.choice(n, size=1500, replace=False).tolist()))
    for _ in range(huge_nets):  # degree > 6144: the arc-parallel arena path
        nets.append(sorted(rng.choice(n, size=7000, replace=False).tolist()))
    noff, pins = csr(nets)
    return n, noff, pins


def compare(ctx, n, noff, pins, nw=None, weight=0, seed=3, sub_rounds=8, sampling=200000, max_passes=5,
            min_frac=0.01):
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins, None, nw)
    comm, st = dh.louvain(edge_weight=weight, seed=seed, sub_rounds=sub_rounds, sampling=sampling,
                          max_passes=max_passes, min_move_fraction=min_frac)
    oh = orc.Hypergraph(n, noff, pins, None, nw)
    g = orc.Graph(oh, weight, ranked=True)
    C, ost = g.louvain(sync=True, seed=seed, sub_rounds=sub_rounds, sampling=sampling, max_passes=max_passes,
                       min_frac=min_frac)
    assert np.array_equal(dh.arrays()["comm"], comm)
    dh.close()
    return comm, C[:n], st, ost


def test_fixture_star_expansion(ctx):
    fx = vectors()["fixture_hypergraph"]
    noff, pins = csr(fx["nets"])
    comm, ref, st, ost = compare(ctx, fx["n"], noff, pins, sub_rounds=1)
    assert comm.tolist() == ref.tolist()
    assert (st["arcs_swept"], st["rounds"], st["levels"], st["moves"]) == \
        (ost["arcs"], ost["rounds"], ost["levels"], ost["moves"])


def test_karate_club_plain_graph_path(ctx):
    n, noff, pins, vw, nw = read_hgr(os.path.join(GOLDEN, vectors()["karate_club"]["instance"]))
    for sub_rounds in (1, 4):
        comm, ref, st, ost = compare(ctx, n, noff, pins, nw=nw, sub_rounds=sub_rounds, max_passes=100,
                                     min_frac=0.0001)
        assert comm.tolist() == ref.tolist()
        assert st["levels"] == ost["levels"] and st["moves"] == ost["moves"]


def test_graph_clustering_of_every_node(ctx):
    """mkp_louvain_graph: the clustering of every node of the Louvain graph as
    run_parallel_louvain returns it (parallel_louvain.h:27), before the
    restriction to the hypernodes: N + E nodes of the star expansion with net
    e at node N + e.  The hypernodes' part equals mkp_louvain's communities,
    and net e carries the oracle's label of its net node (the oracle numbers
    net nodes by locality rank; the rank of a net is read from the arcs of
    its pins, whose incident nets are listed in ascending order)."""
    n, noff, pins = powerlaw(6000, 8000, 300, 23, heavy_nets=1)
    m = len(noff) - 1
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins)
    full, st = dh.louvain_graph(seed=4)
    comm, _ = dh.louvain(seed=4)
    dh.close()
    assert len(full) == n + m
    assert np.array_equal(full[:n], comm)
    g = orc.Graph(orc.Hypergraph(n, noff, pins), 0, ranked=True)
    C, _ = g.louvain(sync=True, seed=4)
    arr = g.arrays()
    net_of_pin = np.repeat(np.arange(m), np.diff(noff.astype(np.int64)))
    order = np.lexsort((net_of_pin, pins))  # by vertex, then by net
    heads = arr["head"][:int(arr["idx"][n])].astype(np.int64)
    rank = np.full(m, -1, np.int64)
    rank[net_of_pin[order]] = heads - n
    assert (rank >= 0).all() and sorted(rank.tolist()) == list(range(m))
    assert np.array_equal(full[n:], C[n + rank])
    assert np.array_equal(full[:n], C[:n])
    # a hypergraph whose nets all have two pins is its own Louvain graph
    kn, knoff, kpins, kvw, knw = read_hgr(os.path.join(GOLDEN, vectors()["karate_club"]["instance"]))
    kh = mtk_amd.Hypergraph(ctx, kn, knoff, kpins, None, knw)
    kfull, _ = kh.louvain_graph(seed=2, sub_rounds=4)
    kcomm, _ = kh.louvain(seed=2, sub_rounds=4)
    kh.close()
    assert len(kfull) == kn and np.array_equal(kfull, kcomm)


@pytest.mark.parametrize("seed,sub_rounds", [(1, 1), (2, 4), (3, 8), (4, 16)])
def test_powerlaw_bit_exact(ctx, seed, sub_rounds):
    n, noff, pins = powerlaw(6000, 9000, 400, seed, heavy_nets=3)
    comm, ref, st, ost = compare(ctx, n, noff, pins, seed=seed, sub_rounds=sub_rounds)
    assert comm.tolist() == ref.tolist()
    assert (st["arcs_swept"], st["rounds"], st["levels"], st["moves"]) == \
        (ost["arcs"], ost["rounds"], ost["levels"], ost["moves"])
    assert st["levels"] >= 2


@pytest.mark.parametrize("seed,sub_rounds", [(5, 1), (6, 8)])
def test_heavy_classes_bit_exact(ctx, seed, sub_rounds):
    """Net nodes of 1,500 and of 7,000 pins beside the small classes: a level
    with a node above 6,144 arcs sends every node above 1,024 arcs to the
    arc-parallel arena path (without one, as in test_powerlaw_bit_exact, they
    take a 1024-thread block and an 8192-slot LDS map each)."""
    n, noff, pins = powerlaw(12000, 15000, 600, seed, heavy_nets=3, huge_nets=2)
    comm, ref, st, ost = compare(ctx, n, noff, pins, seed=seed, sub_rounds=sub_rounds)
    assert comm.tolist() == ref.tolist()
    assert (st["arcs_swept"], st["rounds"], st["levels"], st["moves"]) == \
        (ost["arcs"], ost["rounds"], ost["levels"], ost["moves"])


def test_sampling_threshold(ctx):
    n, noff, pins = powerlaw(3000, 4000, 300, 7, heavy_nets=2)
    comm, ref, st, ost = compare(ctx, n, noff, pins, sampling=12)
    assert comm.tolist() == ref.tolist()


def test_weighted_nets_uniform_weight_function(ctx):
    n, noff, pins = powerlaw(4000, 5000, 100, 9)
    nw = np.random.default_rng(9).integers(1, 5, len(noff) - 1).astype(np.int32)
    comm, ref, st, ost = compare(ctx, n, noff, pins, nw=nw)
    assert comm.tolist() == ref.tolist()


@pytest.mark.parametrize("weight", [1, 2])
def test_fractional_weights_modularity(ctx, weight):
    """non_uniform / degree arc weights are fractional: sums depend on the
    order of atomic additions, so only the quality is compared."""
    n, noff, pins = powerlaw(4000, 2500, 60, 11)
    comm, ref, st, ost = compare(ctx, n, noff, pins, weight=weight)
    oh = orc.Hypergraph(n, noff, pins)
    g = orc.Graph(oh, weight, ranked=True)
    nn = g.sizes()[0]
    # evaluate both hypernode clusterings on the star expansion with every
    # net node a singleton under its own node id (ids >= n, unused by the
    # hypernode communities, which are < n)
    def q(c):
        assert c.max() < n
        full = np.concatenate([c, np.arange(n, nn, dtype=np.int32)]) if nn > n else c
        return g.modularity(full.astype(np.int32))
    assert abs(q(comm) - q(ref)) < 0.02
    nc, nref = len(set(comm.tolist())), len(set(ref.tolist()))
    assert abs(nc - nref) <= 0.2 * nref + 2


def test_group_lane_shapes_and_autotuner_agree(ctx):
    """The lanes per node of the 17..64-arc class (8, 16, 32, 64, or chosen by
    the autotuner) change only the speed: communities and statistics are
    identical, and equal to the oracle's."""
    n, noff, pins = powerlaw(8000, 12000, 200, 17, heavy_nets=1)
    c2 = mtk_amd.Context(0)
    try:
        out = []
        for lanes in (0, 8, 16, 32, 64, 0):
            c2.set_tuning("louvain_group_lanes", lanes)
            dh = mtk_amd.Hypergraph(c2, n, noff, pins)
            out.append(dh.louvain(seed=5))
            dh.close()
        for comm, st in out[1:]:
            assert np.array_equal(comm, out[0][0])
            assert (st["arcs_swept"], st["rounds"], st["levels"], st["moves"]) == \
                (out[0][1]["arcs_swept"], out[0][1]["rounds"], out[0][1]["levels"], out[0][1]["moves"])
        g = orc.Graph(orc.Hypergraph(n, noff, pins), 0, ranked=True)
        C, _ = g.louvain(sync=True, seed=5)
        assert out[0][0].tolist() == C[:n].tolist()
        with pytest.raises(RuntimeError):
            c2.set_tuning("louvain_group_lanes", 24)
    finally:
        c2.close()


def test_fused_class_launch_changes_only_the_speed(ctx):
    """The degree classes up to 1024 arcs of a sub-round in one fused launch
    (the 17..64-arc class in it when small, never, or always) or one launch
    each, with every lane shape of the 17..64-arc class (8 lanes keep their
    own launch): communities and statistics are identical
    and equal to the oracle's.  The hypergraph has nodes in every class,
    129..1024-arc net nodes (a block each) and net nodes above 1024 arcs (the
    arc-parallel path) included."""
    n, noff, pins = powerlaw(20000, 30000, 900, 19, heavy_nets=2, huge_nets=1)
    c2 = mtk_amd.Context(0)
    try:
        out = []
        for fused, lanes in ((1, 0), (0, 0), (2, 0), (3, 8), (3, 16), (3, 32), (3, 64), (0, 64)):
            print("case: fused", fused, "lanes", lanes, flush=True)
            c2.set_tuning("louvain_fused_classes", fused)
            c2.set_tuning("louvain_group_lanes", lanes)
            dh = mtk_amd.Hypergraph(c2, n, noff, pins)
            out.append(dh.louvain(seed=7))
            dh.close()
        for comm, st in out[1:]:
            assert np.array_equal(comm, out[0][0])
            assert (st["arcs_swept"], st["rounds"], st["levels"], st["moves"]) == \
                (out[0][1]["arcs_swept"], out[0][1]["rounds"], out[0][1]["levels"], out[0][1]["moves"])
        g = orc.Graph(orc.Hypergraph(n, noff, pins), 0, ranked=True)
        C, _ = g.louvain(sync=True, seed=7)
        assert out[0][0].tolist() == C[:n].tolist()
        with pytest.raises(RuntimeError):
            c2.set_tuning("louvain_fused_classes", 4)
    finally:
        c2.close()


@pytest.mark.parametrize("state", ["random", "louvain"])
def test_fp64_gains_targets_and_modularity(ctx, state):
    """computeMaxGainCluster on the device for every node of the star
    expansion against the oracle (local_moving_modularity.h:87-129): the best
    gain in FP64 within 1e-12, the same target (or staying, including the
    singleton rule of the synchronous rounds), and metrics::modularity
    (local_moving_modularity.cpp:28-56) within 1e-12."""
    print("case: FP64 gains and modularity, %s clustering" % state, flush=True)
    n, noff, pins = powerlaw(5000, 7000, 300, 21, heavy_nets=2)
    oh = orc.Hypergraph(n, noff, pins)
    g = orc.Graph(oh, orc.Graph.UNIFORM, ranked=True)
    nodes = g.sizes()[0]
    if state == "random":
        rng = np.random.default_rng(3)
        C = rng.integers(0, nodes // 20, nodes).astype(np.int32)
    else:
        C, _ = g.louvain(sync=True, seed=5)
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins)
    tg, gain, q = dh.louvain_evaluate(C)
    size = np.bincount(C, minlength=nodes)
    worst = 0.0
    for u in range(nodes):
        to, og = g.best_cluster(C, u)
        worst = max(worst, abs(og - gain[u]))
        frm = int(C[u])
        if to == frm or (size[frm] == 1 and size[to] == 1 and to > frm):
            assert tg[u] == -1, u
        else:
            assert tg[u] == to, u
    assert worst <= 1e-12, worst
    assert abs(q - g.modularity(C)) <= 1e-12
    dh.close()
