"""Parity of the HIP coarsening (synchronous clustering pass, canonical
contraction, level hierarchy, projection) and of the whole multilevel
pipeline with the oracle.  Clusterings, coarse hypergraphs and final block
ids are integer state and compared bit for bit."""
import os

import numpy as np
import pytest

import mtk_amd
import oracle as orc
from helpers import GOLDEN, csr, read_hgr, vectors

pytestmark = pytest.mark.gpu

V = vectors()


@pytest.fixture(scope="module")
def ctx():
    c = mtk_amd.Context(0)
    yield c
    c.close()


def powerlaw(n, m, max_size, seed, window=512, heavy_nets=0, big_nets=0):
    h = mtk_amd.HostHypergraph.powerlaw(n, m, 2.3, 2, max_size, window=window, local_fraction=0.8, seed=seed)
    a = h.arrays()
    h.close()
    nets = [a["pins"][a["noff"][e]:a["noff"][e + 1]].tolist() for e in range(len(a["noff"]) - 1)]
    rng = np.random.default_rng(seed)
    for _ in range(heavy_nets):  # nets above the ignore threshold
        nets.append(sorted(rng.choice(n, size=1200, replace=False).tolist()))
    for _ in range(big_nets):  # large but rated nets
        nets.append(sorted(rng.choice(n, size=700, replace=False).tolist()))
    noff, pins = csr(nets)
    return n, nets, noff, pins


def both(ctx, n, noff, pins, vw=None, nw=None, comm=None):
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins, vw, nw)
    oh = orc.Hypergraph(n, noff, pins, vw, nw)
    if comm is not None:
        dh.set_communities(comm)
        oh.set_communities(comm)
    return dh, oh


def same_hypergraph(dh, oh):
    a, b = dh.arrays(), oh.arrays()
    for key in ("noff", "pins", "voff", "inc", "vw", "nw", "comm"):
        assert np.array_equal(a[key], b[key]), key
    i = dh.info()
    assert (i.num_vertices, i.num_nets, i.num_pins, i.max_net_size, i.total_weight) == oh.sizes()


# ---------------------------------------------------------------- contraction
def test_contraction_vectors(ctx):
    fx = V["fixture_hypergraph"]
    hc = V["hypergraph_contract"]
    noff, pins = csr(fx["nets"])
    dh = mtk_amd.Hypergraph(ctx, fx["n"], noff, pins)
    for case in hc["cases"]:
        c, cmap = dh.contract(case["cluster"])
        i = c.info()
        assert cmap.tolist() == case["cmap"]
        assert (i.num_vertices, i.num_nets, i.num_pins, i.max_net_size, i.total_weight) == \
            (case["n"], case["m"], case["p"], case["max_esize"], 7)
        a = c.arrays()
        assert a["vw"].tolist() == case["vw"] and a["nw"].tolist() == case["nw"]
        for e, p in enumerate(case["pins"]):
            assert sorted(a["pins"][int(a["noff"][e]):int(a["noff"][e + 1])].tolist()) == sorted(p)
        for u, inc in enumerate(case["inc"]):
            assert sorted(a["inc"][int(a["voff"][u]):int(a["voff"][u + 1])].tolist()) == sorted(inc)
        c.close()
    dh.set_communities(fx["communities"])
    c, _ = dh.contract(hc["communities"]["cluster"])
    assert c.arrays()["comm"].tolist() == hc["communities"]["expected"]
    c.close()
    dh.close()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_contraction_random_clusters(ctx, seed):
    rng = np.random.default_rng(seed)
    n, nets, noff, pins = powerlaw(5000, 7000, 300, seed, heavy_nets=2)
    vw = rng.integers(1, 5, n).astype(np.int32)
    nw = rng.integers(1, 4, len(nets)).astype(np.int32)
    comm = rng.integers(0, 50, n).astype(np.int32)
    dh, oh = both(ctx, n, noff, pins, vw, nw, comm)
    # clusters with many parallel and single-pin nets after contraction
    cluster = (rng.integers(0, n // 8, n) * 8 % n).astype(np.uint32)
    c, cmap = dh.contract(cluster)
    oc, ocmap = oh.contract(cluster)
    assert np.array_equal(cmap, ocmap)
    same_hypergraph(c, oc)
    c.close()
    dh.close()


def test_contraction_identity_and_all_in_one(ctx):
    n, nets, noff, pins = powerlaw(2000, 3000, 100, 9)
    dh, oh = both(ctx, n, noff, pins)
    for cluster in (np.arange(n, dtype=np.uint32), np.full(n, 17, np.uint32)):
        c, cmap = dh.contract(cluster)
        oc, ocmap = oh.contract(cluster)
        assert np.array_equal(cmap, ocmap)
        same_hypergraph(c, oc)
        c.close()
    dh.close()


# ------------------------------------------------------------ clustering pass
@pytest.mark.parametrize("seed,sub_rounds,pass_nr", [(1, 1, 0), (2, 4, 1), (3, 16, 2), (4, 64, 0)])
def test_cluster_pass_bit_exact(ctx, seed, sub_rounds, pass_nr):
    rng = np.random.default_rng(seed)
    n, nets, noff, pins = powerlaw(6000, 9000, 400, seed, heavy_nets=2, big_nets=3)
    vw = rng.integers(1, 4, n).astype(np.int32)
    comm = rng.integers(0, 40, n).astype(np.int32)
    dh, oh = both(ctx, n, noff, pins, vw, None, comm)
    kw = dict(contraction_limit=320, max_node_weight=12, seed=seed, sub_rounds=sub_rounds)
    cl, after, visits = dh.cluster_pass(pass_nr, **kw)
    ocl, oafter, ovisits = oh.cluster_pass(True, pass_nr, with_visits=True, **kw)
    assert np.array_equal(cl, ocl)
    assert (after, visits) == (oafter, ovisits)
    assert len(np.unique(cl)) == after
    dh.close()


@pytest.mark.parametrize("n", [8000, 20000])
@pytest.mark.parametrize("passes", [16, 1, 0])
def test_cluster_pass_heavy_tiers(ctx, passes, n):
    """Proposers with more than 1024 rating keys are rated by one block each
    with LDS maps: on levels of at most 16,384 clusters by cluster id (n =
    8000), otherwise by hash (n = 20000), in one pass when their keys fit and
    else in up to `rating_block_passes` passes over hash classes of the keys;
    the others in the global arena.  Fewer passes (1: only the one-pass
    proposers in LDS; 0: everything in the arena) change no decision."""
    rng = np.random.default_rng(11)
    n, nets, noff, pins = powerlaw(n, n * 5 // 4, 300, 11)
    # 24 nets of 700 pins inside 3,000 vertices: their pins sweep up to ~14
    # of them, so proposers need one to four passes
    for _ in range(24):
        nets.append(sorted((1000 + rng.choice(3000, size=700, replace=False)).tolist()))
    noff, pins = csr(nets)
    vw = rng.integers(1, 3, n).astype(np.int32)
    comm = (np.arange(n) // 4000).astype(np.int32)
    dh, oh = both(ctx, n, noff, pins, vw, None, comm)
    sizes = np.diff(noff).astype(np.int64)
    ub = np.bincount(pins, weights=np.repeat(np.where(sizes < 1000, sizes, 0), sizes), minlength=n)
    assert (ub > 6144).sum() > 100 and ((ub > 4096) & (ub <= 6144)).sum() > 100
    assert ((ub > 1024) & (ub <= 4096)).sum() > 100
    ctx.set_tuning("rating_block_passes", passes)
    try:
        for sub_rounds, pass_nr in ((1, 0), (8, 1)):
            kw = dict(contraction_limit=200, max_node_weight=40, seed=3, sub_rounds=sub_rounds)
            cl, after, visits = dh.cluster_pass(pass_nr, **kw)
            ocl, oafter, ovisits = oh.cluster_pass(True, pass_nr, with_visits=True, **kw)
            assert np.array_equal(cl, ocl), (sub_rounds, pass_nr)
            assert (after, visits) == (oafter, ovisits)
    finally:
        ctx.set_tuning("rating_block_passes", 4)
    dh.close()


def test_cluster_pass_weight_bound_and_sampling(ctx):
    rng = np.random.default_rng(5)
    n, nets, noff, pins = powerlaw(4000, 6000, 200, 5)
    vw = rng.integers(1, 6, n).astype(np.int32)
    nw = rng.integers(1, 5, len(nets)).astype(np.int32)
    dh, oh = both(ctx, n, noff, pins, vw, nw, np.zeros(n, np.int32))
    for kw in (dict(max_node_weight=6, sampling=8), dict(max_node_weight=1000, max_shrink=1.2),
               dict(max_node_weight=8, ignore=5)):
        cl, after, visits = dh.cluster_pass(0, contraction_limit=100, seed=7, **kw)
        ocl, oafter, ovisits = oh.cluster_pass(True, 0, contraction_limit=100, seed=7, with_visits=True, **kw)
        assert np.array_equal(cl, ocl), kw
        assert (after, visits) == (oafter, ovisits)
        w = np.bincount(cl, weights=vw, minlength=n)
        sizes = np.bincount(cl, minlength=n)
        assert np.all((w <= kw["max_node_weight"]) | (sizes <= 1))
    dh.close()


def test_coarsener_chain_vectors(ctx):
    cc = V["coarsener_chain"]
    noff, pins = csr(cc["nets"])
    dh = mtk_amd.Hypergraph(ctx, cc["n"], noff, pins)
    dh.set_communities([u // cc["community_divisor"] for u in range(cc["n"])])
    for seed in range(5):
        h = dh.coarsen(contraction_limit=cc["contraction_limit"], max_node_weight=2 ** 31 - 1, min_shrink=1.0,
                       max_shrink=4.0, seed=seed)
        top = h.level(h.num_levels() - 1)
        i = top.info()
        assert (i.num_vertices, i.num_pins, i.num_nets) == (4, cc["expected_pins"], cc["expected_nets"])
        assert top.arrays()["nw"].tolist() == [cc["expected_net_weight"]] * i.num_nets
        h.close()
    dh.close()


def test_hierarchy_matches_oracle(ctx):
    rng = np.random.default_rng(8)
    n, nets, noff, pins = powerlaw(12000, 16000, 500, 8, heavy_nets=1)
    comm = (np.arange(n) // 300).astype(np.int32)
    dh, oh = both(ctx, n, noff, pins, None, None, comm)
    kw = dict(contraction_limit=320, max_node_weight=int(np.ceil(n / 320)), seed=11)
    h = dh.coarsen(**kw)
    levels, coarsest = oh.coarsen(True, **kw)
    assert h.num_levels() == levels + 1
    same_hypergraph(h.level(h.num_levels() - 1), coarsest)
    assert h.level(h.num_levels() - 1).info().num_vertices < n
    h.close()
    dh.close()


def test_projection(ctx):
    n, nets, noff, pins = powerlaw(6000, 8000, 200, 12)
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins)
    h = dh.coarsen(contraction_limit=1000, max_node_weight=40, seed=1)
    L = h.num_levels()
    assert L >= 2
    k = 4
    coarse = h.level(1)
    cn = coarse.info().num_vertices
    cp = mtk_amd.Partition(coarse, k)
    cpart = (np.arange(cn) % k).astype(np.int32)
    cp.set_parts(cpart)
    fp = mtk_amd.Partition(dh, k)
    h.project(fp, cp, 0)
    _, cmap = dh.contract(dh.cluster_pass(0, contraction_limit=1000, max_node_weight=40, seed=1)[0])
    assert np.array_equal(fp.parts(), cpart[cmap])
    assert fp.km1() == orc.Hypergraph(n, noff, pins).phg_init(k, cpart[cmap])["km1"]
    fp.close(); cp.close(); h.close(); dh.close()


# --------------------------------------------------------------- pipeline
@pytest.mark.parametrize("k,seed", [(2, 1), (8, 2), (32, 3)])
def test_pipeline_bit_exact(ctx, k, seed):
    n, nets, noff, pins = powerlaw(20000, 30000, 600, seed, heavy_nets=2)
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins)
    part, st = dh.partition(k, 0.03, seed)
    opart, okm1, ost = orc.Hypergraph(n, noff, pins).partition(k, 0.03, seed, sync=True)
    assert np.array_equal(part, opart)
    assert st["km1"] == okm1
    assert st["levels"] == ost["levels"] + 1
    assert (st["louvain_arcs"], st["rating_visits"], st["lp_visits"]) == \
        (ost["louvain_arcs"], ost["rating_visits"], ost["lp_visits"])
    perfect = int(np.ceil(n / k))
    assert np.bincount(part, minlength=k).max() <= int(1.03 * perfect)
    dh.close()


def test_pipeline_ibm01(ctx):
    n, noff, pins, vw, nw = read_hgr(os.path.join(GOLDEN, V["lp_invariants"]["instance"]))
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins, vw, nw)
    for k in (2, 4, 8):
        part, st = dh.partition(k, 0.03, 42)
        opart, okm1, _ = orc.Hypergraph(n, noff, pins, vw, nw).partition(k, 0.03, 42, sync=True)
        assert np.array_equal(part, opart)
        assert st["km1"] == okm1
    dh.close()
