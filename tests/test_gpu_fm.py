"""Device FM (multitry localized k-way FM with gain cache and global
rollback, synchronous form) against the oracle's restatement: block ids, km1
and statistics bit for bit, from the same partition."""
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


def instance(n, m, max_size, seed, weighted=False):
    h = mtk_amd.HostHypergraph.powerlaw(n, m, 2.4, 2, max_size, window=max(64, n // 32), local_fraction=0.8,
                                        seed=seed)
    a = h.arrays()
    h.close()
    rng = np.random.default_rng(seed)
    vw = rng.integers(1, 9, n).astype(np.int32) if weighted else None
    nw = rng.integers(1, 4, len(a["noff"]) - 1).astype(np.int32) if weighted else None
    return n, a["noff"], a["pins"], vw, nw


def run_both(ctx, n, noff, pins, vw, nw, k, seed, eps=0.03, lp_first=True, sub_rounds=8):
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins, vw, nw)
    oh = orc.Hypergraph(n, noff, pins, vw, nw)
    w = np.ones(n, np.int64) if vw is None else vw.astype(np.int64)
    total = int(w.sum())
    perfect = -(-total // k)
    maxw = [int((1 + eps) * perfect)] * k
    part = (np.arange(n, dtype=np.int64) * k // n).astype(np.int32)
    if lp_first:
        part, _, _ = oh.lp_refine(k, part, maxw, [perfect] * k, seed=seed, sync=True)
    dp = mtk_amd.Partition(dh, k)
    dp.set_parts(part)
    km1_start = dp.km1()
    st = dp.fm_refine(maxw, seed=seed, sub_rounds=sub_rounds)
    rp, rkm1, rst = oh.fm_refine(k, part, maxw, seed=seed, sub_rounds=sub_rounds)
    assert np.array_equal(dp.parts(), rp), "FM block ids differ from the oracle"
    assert dp.km1() == rkm1
    for key in ("improvement", "moves", "rounds", "searches"):
        assert st[key] == rst[key], key
    assert km1_start - rkm1 == rst["improvement"]
    bw = np.bincount(rp, weights=w, minlength=k)
    assert bw.max() <= maxw[0]
    dp.close()
    dh.close()
    return km1_start, rkm1, rst


@pytest.mark.parametrize("k,seed,sub_rounds", [(2, 1, 8), (3, 2, 1), (8, 3, 8), (8, 4, 3)])
def test_fm_matches_oracle(ctx, k, seed, sub_rounds):
    n, noff, pins, vw, nw = instance(6000, 7000, 120, seed)
    km1_start, km1_end, st = run_both(ctx, n, noff, pins, vw, nw, k, seed, sub_rounds=sub_rounds)
    assert st["improvement"] > 0


@pytest.mark.parametrize("k", [32, 64, 128])
def test_fm_weighted_large_k(ctx, k):
    n, noff, pins, vw, nw = instance(8000, 9000, 200, 7, weighted=True)
    run_both(ctx, n, noff, pins, vw, nw, k, 5)


def test_fm_from_unrefined_partition(ctx):
    n, noff, pins, vw, nw = instance(5000, 6000, 60, 11)
    km1_start, km1_end, st = run_both(ctx, n, noff, pins, vw, nw, 4, 9, lp_first=False)
    assert km1_end < km1_start


@pytest.mark.parametrize("k,seed", [(2, 42), (8, 43)])
def test_pipeline_with_fm_bit_exact(ctx, k, seed):
    """The multilevel pipeline with FM after LP at every level (10 rounds)
    against the oracle's pipeline, and its km1 against the LP-only run."""
    h = mtk_amd.HostHypergraph.powerlaw(20000, 24000, 2.5, 2, 300, window=1024, local_fraction=0.8, seed=seed)
    a = h.arrays()
    n = len(a["vw"])
    dh = mtk_amd.Hypergraph(ctx, host=h)
    part, st = dh.partition(k, 0.03, seed, fm_rounds=10)
    lp_only, lst = dh.partition(k, 0.03, seed)
    opart, okm1, ost = orc.Hypergraph(n, a["noff"], a["pins"], a["vw"], a["nw"]).partition(k, 0.03, seed, sync=True,
                                                                                          fm_rounds=10)
    assert np.array_equal(part, opart) and st["km1"] == okm1
    assert st["fm_improvement"] == ost["fm_improvement"] > 0
    assert st["km1"] <= lst["km1"]
    maxw = int(1.03 * -(-n // k))
    assert np.bincount(part, minlength=k).max() <= maxw
    dh.close()
    h.close()
