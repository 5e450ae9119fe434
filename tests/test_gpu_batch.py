"""The batched form: many hypergraphs refined in the same launches must give
every member exactly the result it gets alone (parts, km1 and statistics),
including members that stop early, large-k members (wave gain kernel) and
batches with more than 1024 blocks (split admission sort)."""
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


def member(n, m, seed, weighted=False, max_size=40):
    return mtk_amd.HostHypergraph.powerlaw(n, m, 2.5, 2, max_size, window=256, local_fraction=0.8,
                                           vertex_weight_min=1, vertex_weight_max=6 if weighted else 1, seed=seed)


def alone(ctx, host, k, part, seed, eps, sub_rounds):
    a = host.arrays()
    n = len(a["vw"])
    dh = mtk_amd.Hypergraph(ctx, n, a["noff"], a["pins"], a["vw"], a["nw"])
    dp = mtk_amd.Partition(dh, k)
    dp.set_parts(part)
    W = int(a["vw"].sum())
    perfect = -(-W // k)
    maxw = int((1.0 + eps) * perfect)
    st = dp.lp_refine([maxw] * k, [perfect] * k, seed=seed, sub_rounds=sub_rounds)
    out = (dp.parts(), dp.km1(), st)
    dp.close()
    dh.close()
    return out, a


def run_batch(ctx, hosts, ks, seeds, eps, sub_rounds, rng):
    parts = [rng.integers(0, k, h.info().num_vertices).astype(np.int32) for h, k in zip(hosts, ks)]
    b = mtk_amd.Batch(ctx, hosts)
    bp = mtk_amd.BatchPartition(b, ks)
    bp.set_parts(np.concatenate(parts))
    stats = bp.lp_refine(seeds, epsilon=eps, sub_rounds=sub_rounds)
    out = bp.parts()
    km1 = bp.km1()
    bp.close()
    b.close()
    return parts, out, km1, stats


@pytest.mark.parametrize("ks,eps,sub_rounds", [([2, 3, 8, 16, 5, 2], 0.03, 4), ([4, 40, 2, 9], 0.03, 2),
                                               ([32] * 40, 0.0, 4)])
def test_batch_lp_equals_members_alone(ctx, ks, eps, sub_rounds):
    rng = np.random.default_rng(len(ks))
    hosts = [member(int(rng.integers(300, 3000)), int(rng.integers(300, 3000)), 100 + g, weighted=g % 2 == 1)
             for g in range(len(ks))]
    seeds = [int(s) for s in rng.integers(1, 1 << 40, len(ks))]
    parts, out, km1, stats = run_batch(ctx, hosts, ks, seeds, eps, sub_rounds, rng)
    off = 0
    for g, (h, k) in enumerate(zip(hosts, ks)):
        n = h.info().num_vertices
        (apart, akm1, ast), _ = alone(ctx, h, k, parts[g], seeds[g], eps, sub_rounds)
        assert np.array_equal(out[off:off + n], apart), g
        assert km1[g] == akm1, g
        for key in ("delta", "moves", "reverts", "visits", "rounds"):
            assert stats[g][key] == ast[key], (g, key)
        off += n
    for h in hosts:
        h.close()


def test_batch_member_matches_oracle(ctx):
    rng = np.random.default_rng(7)
    hosts = [member(2000, 2500, 11), member(1500, 1800, 12, weighted=True)]
    ks, seeds = [4, 6], [5, 6]
    parts, out, km1, stats = run_batch(ctx, hosts, ks, seeds, 0.03, 4, rng)
    a = hosts[1].arrays()
    n0 = hosts[0].info().num_vertices
    n = len(a["vw"])
    W = int(a["vw"].sum())
    perfect = -(-W // 6)
    opart, okm1, ost = orc.Hypergraph(n, a["noff"], a["pins"], a["vw"], a["nw"]).lp_refine(
        6, parts[1], [int(1.03 * perfect)] * 6, [perfect] * 6, seed=6, sync=True)
    assert np.array_equal(out[n0:n0 + n], opart) and km1[1] == okm1
    for h in hosts:
        h.close()
