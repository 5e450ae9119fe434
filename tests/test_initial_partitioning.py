"""Initial partitioning of the coarsest hypergraph (host code shared by the
product pipeline and the oracle).  Regression: with weighted vertices and
blocks only a few vertices heavy, greedy growing used to pick the same
too-heavy seed over and over and grow its bookkeeping until memory ran out."""
import numpy as np
import pytest

import mtk_amd
import oracle as orc


@pytest.mark.timeout(60)
def test_weighted_vertices_many_small_blocks_terminates_balanced():
    n, k, eps = 5444, 128, 0.03
    h = mtk_amd.HostHypergraph.powerlaw(n, n, 2.3, 2, 2000, window=1024, local_fraction=0.8,
                                        vertex_weight_min=1, vertex_weight_max=8, seed=7005)
    a = h.arrays()
    h.close()
    part, km1, _ = orc.Hypergraph(n, a["noff"], a["pins"], a["vw"], a["nw"]).partition(k, eps, 1005, True)
    W = int(a["vw"].sum())
    perfect = -(-W // k)
    bw = np.bincount(part, weights=a["vw"], minlength=k)
    assert part.min() >= 0 and part.max() < k
    assert bw.max() <= int((1.0 + eps) * perfect)
    assert km1 > 0
