"""Ragged and degenerate inputs through the whole pipeline: more blocks than
vertices, no nets at all (with and without degree-zero removal), one net
holding every vertex, only single-pin nets, two vertices.  The device
pipeline must give the oracle pipeline's block ids and km1 bit for bit, with
every block id in range."""
import numpy as np
import pytest

import mtk_amd
import oracle as orc

pytestmark = pytest.mark.gpu

CASES = {
    "k_above_n": (5, [[0, 1], [1, 2, 3], [3, 4]], 8),
    "no_nets": (300, [], 4),
    "no_nets_above_contraction_limit": (1000, [], 2),
    "one_net_with_every_vertex": (200, [list(range(200))], 4),
    "single_pin_nets": (50, [[i] for i in range(50)], 2),
    "two_vertices": (2, [[0, 1]], 2),
}


@pytest.fixture(scope="module")
def ctx():
    c = mtk_amd.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("name", sorted(CASES))
def test_pipeline_edge_case_matches_oracle(ctx, name):
    n, nets, k = CASES[name]
    noff = np.zeros(len(nets) + 1, np.uint64)
    for i, e in enumerate(nets):
        noff[i + 1] = noff[i] + len(e)
    pins = np.array([p for e in nets for p in e], np.uint32)
    print("case", name, "n", n, "nets", len(nets), "k", k, flush=True)
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins)
    part, st = dh.partition(k, 0.03, 1)
    dh.close()
    opart, okm1, ost = orc.Hypergraph(n, noff, pins).partition(k, 0.03, 1, True)
    assert part.min() >= 0 and part.max() < k
    assert np.array_equal(part, opart)
    assert st["km1"] == okm1
    assert st["removed_vertices"] == ost["removed_vertices"]
