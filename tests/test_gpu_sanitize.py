"""Degree-zero vertex and large-net removal on the device
(mkp_hypergraph_reduce) against the oracle's compaction, and the whole
pipeline with removals and the bin-packing restore against the oracle
pipeline, bit for bit."""
import numpy as np
import pytest

import mtk_amd
import oracle as orc

pytestmark = pytest.mark.gpu

REMOVED = 0xFFFFFFFF


@pytest.fixture(scope="module")
def ctx():
    c = mtk_amd.Context(0)
    yield c
    c.close()


def with_isolated(n0, extra, seed, big_net=0):
    """Power-law hypergraph plus `extra` vertices without nets (spread over
    the id range) and optionally one net of `big_net` pins."""
    h = mtk_amd.HostHypergraph.powerlaw(n0, n0, 2.3, 2, 300, window=1024, local_fraction=0.8,
                                        vertex_weight_min=1, vertex_weight_max=5, seed=seed)
    a = h.arrays()
    h.close()
    rng = np.random.default_rng(seed)
    n = n0 + extra
    ids = rng.permutation(n).astype(np.uint32)  # old vertex v becomes ids[v]
    pins = ids[a["pins"]]
    noff = a["noff"].astype(np.uint64)
    nw = a["nw"]
    if big_net:
        big = np.sort(rng.choice(n, big_net, replace=False)).astype(np.uint32)
        pins = np.concatenate([pins, big])
        noff = np.append(noff, noff[-1] + big_net).astype(np.uint64)
        nw = np.append(nw, 3).astype(np.int32)
    vw = np.ones(n, np.int32)
    vw[ids[:n0]] = a["vw"]
    vw[ids[n0:]] = rng.integers(1, 6, extra)
    return n, noff, pins, vw, nw


@pytest.mark.parametrize("climit,thr", [(1000, 50), (10_000, 40), (60_000, 1 << 30)])
def test_reduce_matches_oracle(ctx, climit, thr):
    n, noff, pins, vw, nw = with_isolated(8000, 4000, 17)
    comm = (np.arange(n) % 97).astype(np.int32)
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins, vw, nw)
    dh.set_communities(comm)
    oh = orc.Hypergraph(n, noff, pins, vw, nw)
    oh.set_communities(comm)
    red, vmap, rv, rn = dh.reduce(climit, thr)
    ored, ovmap, orv, orn = oh.reduce(climit, thr)
    assert (rv, rn) == (orv, orn)
    if rv == 0 and rn == 0:
        assert red is None and list(vmap) == list(range(n))
        dh.close()
        return
    assert np.array_equal(vmap, ovmap)
    assert rv == min(int((np.diff(dh.arrays()["voff"]) == 0).sum()), max(n - climit, 0))
    a, o = red.arrays(), ored.arrays()
    for key in ("noff", "pins", "voff", "inc", "vw", "nw", "comm"):
        assert np.array_equal(a[key], o[key]), key
    info = red.info()
    assert info.total_weight == int(o["vw"].sum())
    assert info.max_net_size == int(np.diff(o["noff"]).max())
    red.close()
    dh.close()


@pytest.mark.parametrize("k,big", [(4, 0), (8, 0), (4, 52_000)])
def test_pipeline_with_removals_matches_oracle(ctx, k, big):
    n, noff, pins, vw, nw = with_isolated(60_000 if big else 6000, 20_000 if big else 3000, 23 + k, big)
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins, vw, nw)
    part, st = dh.partition(k, 0.03, 9)
    opart, okm1, ost = orc.Hypergraph(n, noff, pins, vw, nw).partition(k, 0.03, 9, sync=True)
    assert st["removed_vertices"] == ost["removed_vertices"] > 0
    assert st["removed_nets"] == ost["removed_nets"] == (1 if big else 0)
    assert np.array_equal(part, opart)
    assert st["km1"] == okm1
    dh.close()
