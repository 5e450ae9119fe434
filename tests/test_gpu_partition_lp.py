"""Parity of the HIP partition state and label propagation with the oracle.
Integer state is compared bit for bit."""
import numpy as np
import pytest

import mtk_amd
import oracle as orc
from helpers import csr, vectors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = mtk_amd.Context(0)
    yield c
    c.close()


def powerlaw(n, m, max_size, seed, window=256, heavy=0, weighted=False):
    h = mtk_amd.HostHypergraph.powerlaw(n, m, 2.3, 2, max_size, window=window, local_fraction=0.8,
                                        vertex_weight_min=1, vertex_weight_max=8 if weighted else 1, seed=seed)
    a = h.arrays()
    h.close()
    nets = [a["pins"][a["noff"][e]:a["noff"][e + 1]].tolist() for e in range(len(a["noff"]) - 1)]
    rng = np.random.default_rng(seed)
    for v in range(heavy):  # vertices of degree > 512 exercise the wave-per-vertex gain kernel
        others = rng.choice(np.arange(1, n), size=600, replace=False)
        nets += [[v, int(o)] for o in others]
    noff, pins = csr(nets)
    vw = a["vw"] if weighted else None
    nw = rng.integers(1, 4, len(nets)).astype(np.int32) if weighted else None
    return n, noff, pins, vw, nw


def both(ctx, n, noff, pins, vw=None, nw=None):
    return mtk_amd.Hypergraph(ctx, n, noff, pins, vw, nw), orc.Hypergraph(n, noff, pins, vw, nw)


def check_state(dp, ohg, k, part):
    ref = ohg.phg_init(k, part)
    assert np.array_equal(dp.pin_counts(), ref["pin_counts"])
    assert np.array_equal(dp.connectivity(), ref["conn"])
    assert np.array_equal(dp.part_weights(), ref["part_w"])
    assert dp.km1() == ref["km1"]
    assert np.array_equal(dp.border(), ohg.border(k, part))


def test_fixture_partition_state(ctx):
    pv = vectors()["partitioned_hypergraph"]
    fx = vectors()["fixture_hypergraph"]
    noff, pins = csr(fx["nets"])
    dh, oh = both(ctx, fx["n"], noff, pins)
    dp = mtk_amd.Partition(dh, pv["k"])
    dp.set_parts(pv["part"])
    assert dp.pin_counts().tolist() == pv["pin_counts"]
    assert dp.part_weights().tolist() == pv["part_weights"]
    check_state(dp, oh, pv["k"], pv["part"])
    dp.close(); dh.close()


@pytest.mark.parametrize("k", [2, 3, 8, 31, 64, 65, 128, 200])
def test_partition_state_random(ctx, k):
    n, noff, pins, vw, nw = powerlaw(5000, 6000, 300, seed=k, weighted=True)
    dh, oh = both(ctx, n, noff, pins, vw, nw)
    a = dh.arrays()
    ref = oh.arrays()
    assert np.array_equal(a["voff"], ref["voff"]) and np.array_equal(a["inc"], ref["inc"])
    part = np.random.default_rng(k).integers(0, k, n).astype(np.int32)
    dp = mtk_amd.Partition(dh, k)
    dp.set_parts(part)
    check_state(dp, oh, k, part)
    assert np.array_equal(dp.parts(), part)
    dp.close(); dh.close()


def run_lp_pair(ctx, n, noff, pins, vw, nw, k, part, eps, seed, sub_rounds=4, max_w=None):
    dh, oh = both(ctx, n, noff, pins, vw, nw)
    W = int(n if vw is None else np.sum(vw))
    perfect = -(-W // k)
    mw = [max_w if max_w is not None else int((1 + eps) * perfect)] * k
    dp = mtk_amd.Partition(dh, k)
    dp.set_parts(part)
    before = dp.km1()
    st = dp.lp_refine(mw, [perfect] * k, seed=seed, sub_rounds=sub_rounds)
    ref_part, ref_km1, ref_st = oh.lp_refine(k, part, mw, [perfect] * k, seed=seed, sub_rounds=sub_rounds, sync=True)
    got = dp.parts()
    assert np.array_equal(got, ref_part)
    assert dp.km1() == ref_km1 == before + st["delta"]
    for key in ("delta", "moves", "reverts", "visits", "rounds"):
        assert st[key] == ref_st[key], key
    check_state(dp, oh, k, got)
    assert max(dp.part_weights()) <= mw[0] or max(dp.part_weights()) <= max(np.bincount(part, minlength=k))
    dp.close(); dh.close()
    return st


@pytest.mark.parametrize("k", [2, 4, 8, 16, 32, 64, 128])
def test_lp_matches_sync_oracle(ctx, k):
    n, noff, pins, vw, nw = powerlaw(4000, 5000, 200, seed=100 + k, heavy=3)
    part = (np.random.default_rng(k).integers(0, k, n)).astype(np.int32)
    st = run_lp_pair(ctx, n, noff, pins, vw, nw, k, part, 0.03, seed=7)
    assert st["moves"] > 0 and st["delta"] < 0


@pytest.mark.parametrize("sub_rounds", [1, 2, 8])
def test_lp_weighted_and_sub_rounds(ctx, sub_rounds):
    n, noff, pins, vw, nw = powerlaw(3000, 4000, 100, seed=5, heavy=2, weighted=True)
    part = (np.arange(n) * 8 // n).astype(np.int32)
    run_lp_pair(ctx, n, noff, pins, vw, nw, 8, part, 0.03, seed=11, sub_rounds=sub_rounds)


def test_lp_admission_under_tight_balance(ctx):
    """Block capacities so tight that candidate arrivals overflow: the
    deterministic (gain, hash, id) admission path."""
    n, noff, pins, vw, nw = powerlaw(3000, 4000, 60, seed=9, weighted=True)
    k = 4
    rng = np.random.default_rng(3)
    part = rng.integers(0, k, n).astype(np.int32)
    W = int(np.sum(vw))
    maxw = int(np.bincount(part, weights=vw, minlength=k).max()) + 20
    run_lp_pair(ctx, n, noff, pins, vw, nw, k, part, 0.0, seed=2, max_w=maxw)


def test_lp_on_fixture_and_empty(ctx):
    fx = vectors()["fixture_hypergraph"]
    noff, pins = csr(fx["nets"])
    run_lp_pair(ctx, fx["n"], noff, pins, None, None, 2, np.array([1, 0, 0, 0, 0, 1, 1], np.int32), 0.5, seed=1)
    # no nets at all: nothing is a border vertex
    noff0 = np.zeros(1, np.uint64)
    pins0 = np.zeros(0, np.uint32)
    dh = mtk_amd.Hypergraph(ctx, 10, noff0, pins0)
    dp = mtk_amd.Partition(dh, 2)
    dp.set_parts(np.arange(10) % 2)
    assert dp.km1() == 0
    st = dp.lp_refine([6, 6], [5, 5])
    assert st["moves"] == 0 and st["delta"] == 0
    dp.close(); dh.close()
