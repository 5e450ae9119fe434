"""The allocation-free border entry point: the same flags as
mkp_partition_get_border on two calls in a row (the second on a smaller
hypergraph, so stale scratch contents would show), and a partition above the
scratch limit refused before any work."""
import numpy as np
import pytest

import mtk_amd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = mtk_amd.Context(0)
    yield c
    c.close()


def border_pair(ctx, n, m, k, seed):
    h = mtk_amd.HostHypergraph.powerlaw(n, m, 2.3, 2, 300, window=512, local_fraction=0.8, seed=seed)
    dh = mtk_amd.Hypergraph(ctx, host=h)
    p = mtk_amd.Partition(dh, k)
    p.set_parts(np.random.default_rng(seed).integers(0, k, n).astype(np.int32))
    want = p.border()
    got = p.border_noalloc()
    p.close(); dh.close(); h.close()
    return want, got


def test_noalloc_border_equals_border_on_two_calls(ctx):
    print("case: border without allocation, 200k then 30k vertices", flush=True)
    want, got = border_pair(ctx, 200_000, 240_000, 8, seed=5)
    assert np.array_equal(want, got) and want.any() and not want.all()
    want, got = border_pair(ctx, 30_000, 20_000, 4, seed=6)
    assert np.array_equal(want, got) and want.any()


def test_noalloc_border_refuses_partitions_above_the_scratch_limit(ctx):
    print("case: border without allocation, 16,777,217 vertices (refused)", flush=True)
    n = (16 << 20) + 1
    noff = np.array([0, 2], np.uint64)
    pins = np.array([0, n - 1], np.uint32)
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins)
    p = mtk_amd.Partition(dh, 2)
    with pytest.raises(mtk_amd.MkpError, match="exceed the scratch limit"):
        p.border_noalloc()
    assert p.border()[0] == 0  # the allocating entry point still serves it
    p.close(); dh.close()
