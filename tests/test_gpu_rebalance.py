"""The km1 rebalancer called on a device partition (mkp_rebalance: the
partition goes to the host, is rebalanced there and comes back with its
state rebuilt) gives the oracle's result, and the device state (block
weights, km1) matches the moved partition."""
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


@pytest.mark.parametrize("k,seed", [(2, 1), (8, 2), (33, 3)])
def test_device_rebalance_matches_oracle(ctx, k, seed):
    n = 5000
    h = mtk_amd.HostHypergraph.powerlaw(n, n, 2.3, 2, 300, window=1024, local_fraction=0.8,
                                        vertex_weight_min=1, vertex_weight_max=4, seed=seed)
    a = h.arrays()
    rng = np.random.default_rng(seed)
    part = rng.integers(0, k, n).astype(np.int32)
    part[rng.random(n) < 0.4] = k - 1  # block k-1 far over its bound
    if k > 2:
        part[part == 0] = 1  # block 0 starts empty
    W = int(a["vw"].sum())
    maxw = int(1.03 * -(-W // k))
    print("k", k, "seed", seed, flush=True)
    dh = mtk_amd.Hypergraph(ctx, host=h)
    dp = mtk_amd.Partition(dh, k)
    dp.set_parts(part)
    delta, moves = dp.rebalance([maxw] * k)
    opart, odelta, omoves, obal = orc.Hypergraph(n, a["noff"], a["pins"], a["vw"], a["nw"]).rebalance(
        k, part, [maxw] * k)
    assert (delta, moves) == (odelta, omoves) and moves > 0
    assert np.array_equal(dp.parts(), opart)
    ref = orc.Hypergraph(n, a["noff"], a["pins"], a["vw"], a["nw"]).phg_init(k, opart)
    assert dp.km1() == ref["km1"]
    assert np.array_equal(dp.part_weights(), ref["part_w"])
    dp.close(); dh.close(); h.close()
