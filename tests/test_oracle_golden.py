"""Pins the CPU oracle (test infrastructure) against the reference's own test
vectors and against brute-force recomputation.  CPU only."""
import os

import numpy as np
import pytest

import oracle as orc
from helpers import GOLDEN, csr, read_hgr, vectors

V = vectors()


def fixture_hg():
    fx = V["fixture_hypergraph"]
    noff, pins = csr(fx["nets"])
    return orc.Hypergraph(fx["n"], noff, pins)


def random_hg(rng, n, m, max_size, weighted=False):
    nets = []
    for _ in range(m):
        s = int(rng.integers(2, max_size + 1))
        nets.append(sorted(rng.choice(n, size=min(s, n), replace=False).tolist()))
    noff, pins = csr(nets)
    vw = rng.integers(1, 5, n).astype(np.int32) if weighted else None
    nw = rng.integers(1, 4, m).astype(np.int32) if weighted else None
    return orc.Hypergraph(n, noff, pins, vw, nw), nets, vw, nw


def km1_brute(nets, nw, part):
    s = 0
    for i, e in enumerate(nets):
        lam = len({int(part[p]) for p in e})
        s += (lam - 1) * (1 if nw is None else int(nw[i]))
    return s


# ----------------------------------------------------------------- louvain --
def test_louvain_best_cluster_vectors():
    g = orc.Graph(fixture_hg(), orc.Graph.UNIFORM)
    for case in V["louvain_best_cluster"]["cases"]:
        to, _ = g.best_cluster(case["C"], case["u"])
        assert to == case["to"], case


def test_karate_club_reference_semantics():
    kc = V["karate_club"]
    n, noff, pins, vw, nw = read_hgr(os.path.join(GOLDEN, kc["instance"]))
    g = orc.Graph(orc.Hypergraph(n, noff, pins, vw, nw), orc.Graph.UNIFORM)
    assert g.sizes()[0] == 34  # all nets have two pins -> plain graph (graph.cpp:309-317)
    C, _ = g.louvain(sync=False, max_passes=kc["max_passes"], min_frac=kc["min_move_fraction"], threads=1,
                     disable_rand=True)
    assert C.tolist() == kc["expected"]


def test_modularity_gain_matches_brute_force():
    """verifyGain (local_moving_modularity.cpp:133-183): the adjusted gain of
    the chosen move equals the brute-force modularity difference."""
    rng = np.random.default_rng(7)
    for trial in range(20):
        hg, _, _, _ = random_hg(rng, 30, 25, 5, weighted=bool(trial % 2))
        g = orc.Graph(hg, [orc.Graph.UNIFORM, orc.Graph.NON_UNIFORM, orc.Graph.DEGREE][trial % 3])
        n, _, tv, _ = g.sizes()
        a = g.arrays()
        C = rng.integers(0, 6, n).astype(np.int32)
        recip = 1.0 / tv
        for u in rng.choice(n, 5, replace=False):
            u = int(u)
            to, best_gain = g.best_cluster(C, u)
            frm = int(C[u])
            if to == frm:
                continue
            cvol = np.zeros(n)
            np.add.at(cvol, C, a["vol"])
            arcs = range(int(a["idx"][u]), int(a["idx"][u + 1]))
            w_from = sum(a["w"][j] for j in arcs if C[a["head"][j]] == frm)
            volU = a["vol"][u]
            adjusted = 2.0 * recip * (best_gain - w_from + recip * volU * (cvol[frm] - volU))
            q0 = g.modularity(C)
            C2 = C.copy()
            C2[u] = to
            q1 = g.modularity(C2)
            assert abs((q0 + adjusted) - q1) < 1e-10


# ------------------------------------------------------------------- graph --
def test_star_expansion_vectors():
    gv = V["graph"]
    hg = fixture_hg()
    g = orc.Graph(hg, orc.Graph.UNIFORM)
    n, arcs, tv, md = g.sizes()
    assert (n, arcs, md, tv) == (gv["num_nodes"], gv["num_arcs"], gv["max_degree"], gv["total_volume_uniform"])
    a = g.arrays()
    for u, vol in gv["volumes_uniform"].items():
        assert a["vol"][int(u)] == vol
    assert np.diff(a["idx"]).tolist() == gv["degrees"]
    gn = orc.Graph(hg, orc.Graph.NON_UNIFORM).arrays()
    for u, vol in gv["volumes_non_uniform"].items():
        assert gn["vol"][int(u)] == pytest.approx(vol, abs=1e-15)
    kinds = {"uniform": orc.Graph.UNIFORM, "non_uniform": orc.Graph.NON_UNIFORM, "degree": orc.Graph.DEGREE}
    for case in gv["arcs"]:
        ga = orc.Graph(hg, kinds[case["weight"]]).arrays()
        s, e = int(ga["idx"][case["u"]]), int(ga["idx"][case["u"] + 1])
        assert ga["head"][s:e].tolist() == case["heads"]
        assert ga["w"][s:e].tolist() == case["w"]


def test_graph_contract_vectors():
    g = orc.Graph(fixture_hg(), orc.Graph.UNIFORM)
    for case in V["graph_contract"]["cases"]:
        cg, _ = g.contract(case["C"])
        n, _, tv, md = cg.sizes()
        assert tv == 24 and md == case["max_degree"] and n == len(case["volumes"])
        a = cg.arrays()
        assert a["vol"].tolist() == case["volumes"]
        for u, (heads, ws) in enumerate(case["arcs"]):
            s, e = int(a["idx"][u]), int(a["idx"][u + 1])
            assert a["head"][s:e].tolist() == heads
            assert a["w"][s:e].tolist() == ws
    tw = V["graph_contract"]["twice"]
    cg, _ = g.contract(tw["C1"])
    ccg, _ = cg.contract(tw["C2"])
    assert ccg.arrays()["vol"].tolist() == tw["volumes"]


# ------------------------------------------------------------ contraction --
def test_hypergraph_contract_vectors():
    hc = V["hypergraph_contract"]
    hg = fixture_hg()
    for case in hc["cases"]:
        c, cmap = hg.contract(case["cluster"])
        n, m, p, me, tw = c.sizes()
        assert cmap.tolist() == case["cmap"]
        assert (n, m, p, me, tw) == (case["n"], case["m"], case["p"], case["max_esize"], 7)
        a = c.arrays()
        assert a["vw"].tolist() == case["vw"] and a["nw"].tolist() == case["nw"]
        for e, pins in enumerate(case["pins"]):
            assert sorted(a["pins"][int(a["noff"][e]):int(a["noff"][e + 1])].tolist()) == sorted(pins)
        for u, inc in enumerate(case["inc"]):
            assert sorted(a["inc"][int(a["voff"][u]):int(a["voff"][u + 1])].tolist()) == sorted(inc)
    hg.set_communities(V["fixture_hypergraph"]["communities"])
    c, _ = hg.contract(hc["communities"]["cluster"])
    assert c.arrays()["comm"].tolist() == hc["communities"]["expected"]


def contract_brute(nets, nw, vw, cluster):
    reps = sorted(set(int(x) for x in cluster))
    rank = {r: i for i, r in enumerate(reps)}
    cmap = [rank[int(c)] for c in cluster]
    cvw = [0] * len(reps)
    for u, c in enumerate(cmap):
        cvw[c] += int(vw[u])
    out = {}
    order = []
    for e, net in enumerate(nets):
        key = tuple(sorted({cmap[p] for p in net}))
        if len(key) < 2:
            continue
        if key not in out:
            out[key] = 0
            order.append(key)
        out[key] += int(nw[e])
    return cmap, cvw, [list(k) for k in order], [out[k] for k in order]


def test_contraction_matches_brute_force():
    rng = np.random.default_rng(3)
    for _ in range(10):
        n, m = 40, 60
        hg, nets, vw, nw = random_hg(rng, n, m, 6, weighted=True)
        cluster = rng.integers(0, 15, n)
        cluster = np.array([c if c < n else 0 for c in cluster], dtype=np.uint32)
        c, cmap = hg.contract(cluster)
        bmap, bvw, bnets, bnw = contract_brute(nets, nw, vw, cluster)
        a = c.arrays()
        assert cmap.tolist() == bmap
        assert a["vw"].tolist() == bvw
        got = [a["pins"][int(a["noff"][e]):int(a["noff"][e + 1])].tolist() for e in range(len(a["nw"]))]
        assert got == bnets  # canonical: ascending pins, nets in first-occurrence order
        assert a["nw"].tolist() == bnw
        for u in range(len(bvw)):  # incident nets ascending
            inc = a["inc"][int(a["voff"][u]):int(a["voff"][u + 1])].tolist()
            assert inc == sorted(e for e, p in enumerate(got) if u in p)


def test_coarsener_chain_reference_semantics():
    cc = V["coarsener_chain"]
    noff, pins = csr(cc["nets"])
    hg = orc.Hypergraph(cc["n"], noff, pins)
    hg.set_communities([u // cc["community_divisor"] for u in range(cc["n"])])
    levels, coarsest = hg.coarsen(sync=False, contraction_limit=cc["contraction_limit"],
                                  max_node_weight=2 ** 31 - 1, min_shrink=1.0, max_shrink=4.0,
                                  disable_rand=True, prefer_unmatched=False)
    n, m, p, _, _ = coarsest.sizes()
    assert (p, m) == (cc["expected_pins"], cc["expected_nets"])
    assert coarsest.arrays()["nw"].tolist() == [cc["expected_net_weight"]] * m


def test_coarsener_chain_sync_semantics_reaches_communities():
    cc = V["coarsener_chain"]
    noff, pins = csr(cc["nets"])
    hg = orc.Hypergraph(cc["n"], noff, pins)
    hg.set_communities([u // cc["community_divisor"] for u in range(cc["n"])])
    for seed in range(5):
        levels, coarsest = hg.coarsen(sync=True, contraction_limit=cc["contraction_limit"],
                                      max_node_weight=2 ** 31 - 1, min_shrink=1.0, max_shrink=4.0, seed=seed)
        n, m, p, _, _ = coarsest.sizes()
        assert (n, p, m) == (4, cc["expected_pins"], cc["expected_nets"])
        assert coarsest.arrays()["nw"].tolist() == [cc["expected_net_weight"]] * m


def test_clustering_respects_communities_and_weight():
    rng = np.random.default_rng(11)
    for sync in (False, True):
        hg, nets, vw, nw = random_hg(rng, 200, 300, 6, weighted=True)
        comm = rng.integers(0, 4, 200).astype(np.int32)
        hg.set_communities(comm)
        cl, after = hg.cluster_pass(sync, 0, contraction_limit=10, max_node_weight=9, seed=3)
        assert len(set(cl.tolist())) == after
        w = {}
        for u, c in enumerate(cl):
            assert cl[c] == c  # representatives point to themselves
            assert comm[u] == comm[c]
            w[int(c)] = w.get(int(c), 0) + int(vw[u])
        assert all(x <= 9 or cl.tolist().count(c) == 1 for c, x in w.items())


# --------------------------------------------------------- partition state --
def test_partitioned_hypergraph_vectors():
    pv = V["partitioned_hypergraph"]
    hg = fixture_hg()
    st = hg.phg_init(pv["k"], pv["part"])
    assert st["part_w"].tolist() == pv["part_weights"]
    assert st["pin_counts"].tolist() == pv["pin_counts"]
    for case in pv["move_cases"]:
        part = list(pv["part"])
        for u, frm, to in case["moves"]:
            assert part[u] == frm
            part[u] = to
        st = hg.phg_init(pv["k"], part)
        if "part_weights" in case:
            assert st["part_w"].tolist() == case["part_weights"]
        if "pin_counts" in case:
            assert st["pin_counts"].tolist() == case["pin_counts"]


def test_pin_counts_connectivity_km1_brute_force():
    rng = np.random.default_rng(5)
    for k in (2, 3, 8, 64, 65, 128):
        hg, nets, vw, nw = random_hg(rng, 100, 150, 12, weighted=True)
        part = rng.integers(0, k, 100).astype(np.int32)
        st = hg.phg_init(k, part)
        for e, net in enumerate(nets):
            cnt = np.bincount(part[net], minlength=k)
            assert st["pin_counts"][e].tolist() == cnt.tolist()
            bits = [(int(st["conn"][e][b // 64]) >> (b % 64)) & 1 for b in range(k)]
            assert bits == (cnt > 0).astype(int).tolist()
        assert st["km1"] == km1_brute(nets, nw, part)
        assert st["part_w"].tolist() == np.bincount(part, weights=vw, minlength=k).astype(int).tolist()


def test_gain_policy_vectors():
    hg = fixture_hg()
    for case in V["gain_policy_km1"]["cases"]:
        k = case["k"]
        to, gain = hg.max_gain_move(k, case["part"], case["u"], [2 ** 40] * k)
        assert (to, gain) == (case["to"], case["gain"]), case
        if "delta" in case:
            part = list(case["part"])
            before = hg.phg_init(k, part)["km1"]
            part[case["u"]] = to
            assert hg.phg_init(k, part)["km1"] - before == case["delta"]


def test_km1_gains_brute_force():
    """verifyAllKm1GainValues (partitioned_hypergraph_test.cc:120-134)."""
    rng = np.random.default_rng(9)
    for k in (2, 4, 7):
        hg, nets, vw, nw = random_hg(rng, 40, 50, 6, weighted=True)
        part = rng.integers(0, k, 40).astype(np.int32)
        base = km1_brute(nets, nw, part)
        for u in range(40):
            to, gain = hg.max_gain_move(k, part, u, [2 ** 40] * k)
            best = min([0] + [km1_brute(nets, nw, np.where(np.arange(40) == u, t, part)) - base
                              for t in range(k) if t != part[u]])
            assert gain == best
            if to != part[u]:
                p2 = part.copy()
                p2[u] = to
                assert km1_brute(nets, nw, p2) - base == gain


# ----------------------------------------------------------------------- LP --
@pytest.mark.parametrize("sync", [False, True])
@pytest.mark.parametrize("k", [2, 4, 8])
def test_lp_invariants_on_ibm01(sync, k):
    """label_propagation_refiner_test.cc:133-153: never worsens, tracked
    delta equals recomputation, imbalance stays within eps + 0.05."""
    lv = V["lp_invariants"]
    n, noff, pins, vw, nw = read_hgr(os.path.join(GOLDEN, lv["instance"]))
    hg = orc.Hypergraph(n, noff, pins, vw, nw)
    W = int(vw.sum())
    perfect = -(-W // k)
    maxw = int((1 + lv["epsilon"]) * perfect)
    part = (np.arange(n) * k // n).astype(np.int32)
    before = hg.phg_init(k, part)["km1"]
    out, km1, st = hg.lp_refine(k, part, [maxw] * k, [perfect] * k, seed=1, sync=sync)
    assert km1 <= before
    assert km1 == before + st["delta"]
    assert hg.phg_init(k, out)["km1"] == km1
    pw = np.bincount(out, weights=vw, minlength=k)
    assert pw.max() / perfect - 1.0 <= lv["epsilon"] + 0.05
    assert pw.max() <= maxw


def test_lp_sync_is_deterministic_and_seeded():
    rng = np.random.default_rng(1)
    hg, nets, vw, nw = random_hg(rng, 500, 700, 8)
    part = rng.integers(0, 4, 500).astype(np.int32)
    a = hg.lp_refine(4, part, [140] * 4, [125] * 4, seed=3, sync=True)
    b = hg.lp_refine(4, part, [140] * 4, [125] * 4, seed=3, sync=True)
    assert a[0].tolist() == b[0].tolist() and a[1] == b[1]
    assert a[1] < hg.phg_init(4, part)["km1"]


def test_louvain_sync_improves_modularity():
    rng = np.random.default_rng(2)
    hg, *_ = random_hg(rng, 300, 400, 5)
    g = orc.Graph(hg, orc.Graph.UNIFORM)
    n = g.sizes()[0]
    C, st = g.louvain(sync=True, seed=5)
    assert g.modularity(C) > g.modularity(np.arange(n, dtype=np.int32)) + 0.1
    C2, _ = g.louvain(sync=True, seed=5)
    assert C.tolist() == C2.tolist()


def test_ranked_star_expansion_is_a_relabelling():
    """The locality-ranked star expansion is the star expansion with the net
    nodes permuted, in non-decreasing order of the lower median of each net's
    first nine pins (ties by net id)."""
    rng = np.random.default_rng(21)
    hg, nets, vw, nw = random_hg(rng, 300, 500, 14, weighted=True)
    n, m = 300, len(nets)
    for kind in (orc.Graph.UNIFORM, orc.Graph.DEGREE):
        a = orc.Graph(hg, kind).arrays()
        b = orc.Graph(hg, kind, ranked=True).arrays()
        # the permutation read off the hypernode arcs
        perm = {}
        for u in range(n):
            ha = a["head"][a["idx"][u]:a["idx"][u + 1]]
            hb = b["head"][b["idx"][u]:b["idx"][u + 1]]
            assert np.array_equal(a["w"][a["idx"][u]:a["idx"][u + 1]], b["w"][b["idx"][u]:b["idx"][u + 1]])
            for x, y in zip(ha.tolist(), hb.tolist()):
                assert perm.setdefault(x, y) == y
        assert sorted(perm) == list(range(n, n + m)) and sorted(perm.values()) == list(range(n, n + m))
        for x, y in perm.items():
            assert a["vol"][x] == b["vol"][y]
            assert np.array_equal(a["head"][a["idx"][x]:a["idx"][x + 1]], b["head"][b["idx"][y]:b["idx"][y + 1]])
        keys = []
        for e in range(m):
            p = sorted(nets[e][:9])
            keys.append((p[(len(p) - 1) // 2], e))
        order = [e for _, e in sorted(keys)]
        assert [perm[n + e] - n for e in order] == list(range(m))


# ------------------------------------------------------------ border nodes --
def test_border_node_vectors():
    """isBorderNode and numIncidentCutHyperedges after the moves of
    partitioned_hypergraph_test.cc:373-466 (fixture partition, k = 3)."""
    bv = V["border_nodes"]
    fx = V["fixture_hypergraph"]
    hg = fixture_hg()
    k = bv["k"]
    for case in bv["cases"]:
        part = list(bv["part"])
        for u, frm, to in case["moves"]:
            assert part[u] == frm
            part[u] = to
        assert hg.border(k, part).tolist() == case["border"], case["_line"]
        st = hg.phg_init(k, part)
        lam = [bin(int(st["conn"][e][0])).count("1") for e in range(len(fx["nets"]))]
        cut_incident = [sum(1 for e, net in enumerate(fx["nets"]) if u in net and lam[e] > 1)
                        for u in range(fx["n"])]
        assert cut_incident == case["cut_incident"], case["_line"]


def test_border_brute_force():
    """A vertex is a border vertex iff some incident net spans more than one
    block, unless its degree exceeds 100,000 (partitioned_hypergraph.h:67,
    605-620): checked on random partitions and on a hub of degree 100,001."""
    rng = np.random.default_rng(17)
    for k in (2, 3, 8, 65):
        hg, nets, _, _ = random_hg(rng, 120, 90, 8, weighted=True)
        part = rng.integers(0, k, 120).astype(np.int32)
        got = hg.border(k, part)
        for u in range(120):
            want = any(u in net and len({int(part[p]) for p in net}) > 1 for net in nets)
            assert bool(got[u]) == want, (k, u)
    # hub vertex 0 in 100,001 two-pin nets, all cut: never a border vertex;
    # vertex 1 (degree 100,000 through the hub's nets) still is one
    n = 100_002
    nets = [[0, 1 + (i % (n - 1))] for i in range(100_001)]
    noff, pins = csr(nets)
    hub = orc.Hypergraph(n, noff, pins)
    part = np.ones(n, np.int32)
    part[0] = 0
    b = hub.border(2, part)
    assert b[0] == 0 and b[1] == 1 and b[2] == 1


# ----------------------------------------------------- synchronous rating --
def test_sync_rating_brute_force():
    """The rating step of a synchronous clustering sub-round against a
    brute-force restatement in exact rationals: for every proposer u the score
    of a cluster is the sum over u's nets e (|e| < ignore) holding a member of
    it of w(e) / (distinct clusters of e - 1), distinct counted at least 2
    (multilevel_vertex_pair_rater.h:109-199, HeavyEdgeScore
    rating_score_policy.h:29-35); candidates are other clusters of u's
    community within the node weight limit; the target has the largest score,
    an unmatched cluster among equal scores (rating_acceptance_policy.h:56-72),
    and its 36.28 fixed-point rating is the score up to the truncation of each
    net's term."""
    from fractions import Fraction
    rng = np.random.default_rng(23)
    for trial in range(6):
        n, m = 150, 170
        nets = []
        for _ in range(m):
            s = int(rng.integers(2, 12))
            c = int(rng.integers(0, n))
            nets.append(sorted({(c + int(x)) % n for x in rng.integers(-10, 11, s)} | {c}))
        nets = [e for e in nets if len(e) >= 2]
        noff, pins = csr(nets)
        vw = rng.integers(1, 4, n).astype(np.int32)
        nw = rng.integers(1, 5, len(nets)).astype(np.int32)
        hg = orc.Hypergraph(n, noff, pins, vw, nw)
        comm = (np.arange(n) // 40).astype(np.int32)
        hg.set_communities(comm)
        # a state: some clusters formed, representatives point to themselves
        cid = np.arange(n, dtype=np.uint32)
        for u in rng.choice(n, 40, replace=False):
            v = int(rng.integers(max(0, u - 5), min(n, u + 6)))
            if comm[v] == comm[u] and cid[v] == v and cid[u] == u and u != v:
                cid[u] = v
        cw = np.zeros(n, np.int64)
        for u in range(n):
            cw[cid[u]] += vw[u]
        size = np.bincount(cid, minlength=n)
        matched = (size[cid] > 1).astype(np.uint8)
        proposers = np.array([u for u in range(n) if not matched[u]], np.uint32)
        maxw = int(rng.integers(4, 9))
        ignore = 9 if trial % 2 else 1000
        tgt, rat = hg.rate_sync(trial, 3, cid, cw, matched, proposers, maxw, ignore=ignore, seed=trial)
        incident = [[e for e, net in enumerate(nets) if u in net] for u in range(n)]
        for u, t, r in zip(proposers.tolist(), tgt.tolist(), rat.tolist()):
            score, terms = {}, {}
            for e in incident[u]:
                if len(nets[e]) >= ignore:
                    continue
                reps = {int(cid[p]) for p in nets[e]}
                d = max(len(reps), 2)
                for c in reps:
                    score[c] = score.get(c, Fraction(0)) + Fraction(int(nw[e]), d - 1)
                    terms[c] = terms.get(c, 0) + 1
            cand = {c: x for c, x in score.items()
                    if c != u and cw[u] + cw[c] <= maxw and comm[c] == comm[u]}
            if not cand:
                assert t == 0xFFFFFFFF, u
                continue
            best = max(cand.values())
            assert t in cand and cand[t] == best, (u, t)
            exact = best * (1 << 28)
            assert exact - terms[t] < r <= exact, (u, r, float(exact))
            tied = [c for c, x in cand.items() if x == best]
            if any(not matched[c] for c in tied):
                assert not matched[t], (u, t)


# ------------------------------------------ synchronous local moving rounds --
def test_louvain_sync_sub_rounds_brute_force():
    """Every sub-round of synchronous local moving on a star expansion,
    replayed from the recorded trace: its nodes are whole tiles of 64
    consecutive ids (each round covers every node once), and its moves are
    exactly the nodes whose best cluster on the state at the sub-round start
    (computeMaxGainCluster, local_moving_modularity.h:87-129, volumes
    recomputed from that state) differs from their own, except a singleton
    moving to a larger singleton; every sampled move alone raises the
    brute-force modularity (local_moving_modularity.cpp:28-56)."""
    h = orc.Hypergraph(*_powerlaw_arrays(1500, 1500, seed=3))
    g = orc.Graph(h, orc.Graph.UNIFORM, ranked=True)
    n = g.sizes()[0]
    final, subs = g.louvain_level_sync_trace(seed=11, sub_rounds=8)
    assert len(subs) >= 8
    C = np.arange(n, dtype=np.int32)
    rng = np.random.default_rng(4)
    for i, (members, mu, mt) in enumerate(subs):
        if i % 8 == 0:
            round_nodes = []
        round_nodes.extend(members.tolist())
        tiles = set((members >> 6).tolist())
        assert sum(min(64, n - (t << 6)) for t in tiles) == len(members)
        csize = np.bincount(C, minlength=n)
        expect = {}
        for u in members.tolist():
            to, _ = g.best_cluster(C, u)
            frm = int(C[u])
            if to != frm and not (csize[frm] == 1 and csize[to] == 1 and to > frm):
                expect[u] = to
        assert dict(zip(mu.tolist(), mt.tolist())) == expect, i
        q0 = g.modularity(C)
        for j in rng.choice(len(mu), min(5, len(mu)), replace=False) if len(mu) else []:
            C2 = C.copy()
            C2[mu[j]] = mt[j]
            assert g.modularity(C2) > q0
        C[mu] = mt
        if i % 8 == 7:
            assert sorted(round_nodes) == list(range(n))
    assert np.array_equal(C, final)


def _powerlaw_arrays(n, m, seed):
    rng = np.random.default_rng(seed)
    nets = []
    for _ in range(m):
        s = int(min(40, 2 + rng.zipf(2.3)))
        c = int(rng.integers(0, n))
        nets.append(sorted({(c + int(x)) % n for x in rng.integers(-60, 61, s)} | {c}))
    nets = [e for e in nets if len(e) >= 2]
    noff, pins = csr(nets)
    return n, noff, pins
