"""Row-net model of a synthetic sparse matrix (rows = weighted vertices,
columns = nets, banded nonzeros): large-k partition state, LP at k = 64,
Louvain and the whole pipeline, each bit-exact against the oracle."""
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


@pytest.fixture(scope="module")
def rownet():
    h = mtk_amd.HostHypergraph.rownet(30000, 30000, 8.0, 600, 3000, 0.7, seed=5)
    a = h.arrays()
    h.close()
    return a


def test_rownet_generator_shape(rownet):
    a = rownet
    n = len(a["vw"])
    sizes = np.diff(a["noff"].astype(np.int64))
    assert sizes.min() >= 2  # single-pin columns dropped, as the hMetis reader does
    assert a["vw"].min() >= 1 and a["vw"].sum() > n  # vertex weight = row nnz
    assert a["pins"].max() < n


@pytest.mark.parametrize("k", [64, 128])
def test_rownet_partition_state_and_lp(ctx, rownet, k):
    a = rownet
    n = len(a["vw"])
    dh = mtk_amd.Hypergraph(ctx, n, a["noff"], a["pins"], a["vw"], a["nw"])
    oh = orc.Hypergraph(n, a["noff"], a["pins"], a["vw"], a["nw"])
    part = (np.arange(n, dtype=np.int64) * k // n).astype(np.int32)  # banded start: blocks of rows
    W = int(a["vw"].sum())
    perfect = -(-W // k)
    maxw = [int(1.03 * perfect)] * k
    dp = mtk_amd.Partition(dh, k)
    dp.set_parts(part)
    ref = oh.phg_init(k, part)
    assert np.array_equal(dp.pin_counts(), ref["pin_counts"])
    assert np.array_equal(dp.connectivity(), ref["conn"])
    assert dp.km1() == ref["km1"]
    st = dp.lp_refine(maxw, [perfect] * k, seed=3)
    opart, okm1, ost = oh.lp_refine(k, part, maxw, [perfect] * k, seed=3, sync=True)
    assert np.array_equal(dp.parts(), opart)
    assert dp.km1() == okm1 and st["delta"] == ost["delta"]
    assert (st["moves"], st["reverts"], st["visits"]) == (ost["moves"], ost["reverts"], ost["visits"])
    assert np.bincount(dp.parts(), weights=a["vw"], minlength=k).max() <= max(maxw[0], ref["part_w"].max())
    dp.close()
    dh.close()


def test_rownet_louvain_and_pipeline(ctx, rownet):
    a = rownet
    n = len(a["vw"])
    dh = mtk_amd.Hypergraph(ctx, n, a["noff"], a["pins"], a["vw"], a["nw"])
    oh = orc.Hypergraph(n, a["noff"], a["pins"], a["vw"], a["nw"])
    comm, st = dh.louvain(edge_weight=0, seed=9)
    C, ost = orc.Graph(oh, orc.Graph.UNIFORM, ranked=True).louvain(sync=True, seed=9)
    assert np.array_equal(comm, C[:n])
    part, pst = dh.partition(64, 0.03, 9)
    opart, okm1, _ = orc.Hypergraph(n, a["noff"], a["pins"], a["vw"], a["nw"]).partition(64, 0.03, 9, sync=True)
    assert np.array_equal(part, opart) and pst["km1"] == okm1
    perfect = -(-int(a["vw"].sum()) // 64)
    assert np.bincount(part, weights=a["vw"], minlength=64).max() <= int(1.03 * perfect)
    dh.close()
