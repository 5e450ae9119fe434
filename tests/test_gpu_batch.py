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


def arrays_member(n, m, max_size, seed, heavy_nets=0, exponent=2.3, heavy_size=1300):
    h = mtk_amd.HostHypergraph.powerlaw(n, m, exponent, 2, max_size, window=256, local_fraction=0.8, seed=seed)
    a = h.arrays()
    h.close()
    nets = [a["pins"][a["noff"][e]:a["noff"][e + 1]] for e in range(len(a["noff"]) - 1)]
    rng = np.random.default_rng(seed)
    # a net node of degree > 1024 (a 1024-thread block each), > 6144 the
    # arc-parallel arena path
    for _ in range(heavy_nets):
        nets.append(np.sort(rng.choice(n, size=heavy_size, replace=False)).astype(np.uint32))
    noff = np.zeros(len(nets) + 1, np.uint64)
    noff[1:] = np.cumsum([len(x) for x in nets])
    return mtk_amd.HostHypergraph.from_arrays(n, noff, np.concatenate(nets))


def louvain_alone(ctx, host, seed, **kw):
    dh = mtk_amd.Hypergraph(ctx, host=host)
    comm, st = dh.louvain(seed=seed, **kw)
    dh.close()
    return comm, st


@pytest.mark.parametrize("sub_rounds,count", [(8, 7), (1, 3), (4, 24)])
def test_batch_louvain_equals_members_alone(ctx, sub_rounds, count):
    rng = np.random.default_rng(sub_rounds * 100 + count)
    hosts = []
    for g in range(count):
        n = int(rng.integers(200, 4000))
        hosts.append(arrays_member(n, int(1.2 * n), int(rng.choice([16, 64, 300])), 500 + g,
                                   heavy_nets=1 if g % 3 == 1 and n > 1300 else 0))
    seeds = [int(s) for s in rng.integers(1, 1 << 40, count)]
    b = mtk_amd.Batch(ctx, hosts)
    comm, stats = b.louvain(seeds=seeds, sub_rounds=sub_rounds)
    dh = b.hypergraph()
    assert np.array_equal(dh.arrays()["comm"], comm)  # resident member-local ids
    off = 0
    for g, h in enumerate(hosts):
        n = h.info().num_vertices
        ac, ast = louvain_alone(ctx, h, seeds[g], sub_rounds=sub_rounds)
        assert np.array_equal(comm[off:off + n], ac), g
        for key in ("arcs_swept", "rounds", "levels", "moves"):
            assert stats[g][key] == ast[key], (g, key)
        off += n
    assert max(s["levels"] for s in stats) >= 2
    b.close()
    for h in hosts:
        h.close()


def test_batch_louvain_member_matches_oracle(ctx):
    hosts = [arrays_member(3000, 3600, 300, 31, heavy_nets=1), arrays_member(1000, 1200, 40, 32),
             arrays_member(9000, 10000, 300, 33, heavy_nets=2, heavy_size=7000)]
    b = mtk_amd.Batch(ctx, hosts)
    comm, stats = b.louvain(seeds=[9, 10, 11])
    off = 0
    for i, (h, seed) in enumerate(zip(hosts, [9, 10, 11])):
        a = h.arrays()
        n = a["n"]
        g = orc.Graph(orc.Hypergraph(n, a["noff"], a["pins"]), 0, ranked=True)
        C, ost = g.louvain(sync=True, seed=seed, sub_rounds=8)
        assert comm[off:off + n].tolist() == C[:n].tolist(), i
        assert (stats[i]["arcs_swept"], stats[i]["rounds"], stats[i]["levels"], stats[i]["moves"]) == \
            (ost["arcs"], ost["rounds"], ost["levels"], ost["moves"]), i
        off += n
    b.close()
    for h in hosts:
        h.close()


def test_batch_louvain_rejects_mixed_weighting(ctx):
    # |E|/|V| < 0.75 resolves the hybrid weighting to "degree", >= 0.75 to "uniform"
    hosts = [arrays_member(1000, 500, 16, 41), arrays_member(1000, 1200, 16, 42)]
    b = mtk_amd.Batch(ctx, hosts)
    with pytest.raises(RuntimeError, match="edge weighting"):
        b.louvain(seeds=[1, 2])
    comm, stats = b.louvain(seeds=[1, 2], edge_weight=0)  # explicit weighting is accepted
    assert len(stats) == 2
    b.close()
    for h in hosts:
        h.close()
