"""The served form: requests of uneven size run concurrently in the request
loop, and each request's partition is the one it gets alone (block ids and
km1 bit for bit), within the (1 + eps) balance bound; admission respects the
HBM budget and the byte estimate bounds what a request holds."""
import time

import numpy as np
import pytest

import mtk_amd

pytestmark = pytest.mark.gpu

KS = (2, 8, 32, 128)


@pytest.fixture(scope="module")
def ctx():
    c = mtk_amd.Context(0)
    yield c
    c.close()


def requests(count, seed, pmin=3_000, pmax=300_000):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        p = float(np.exp(rng.uniform(np.log(pmin), np.log(pmax))))
        n = max(64, int(p / 6))
        h = mtk_amd.HostHypergraph.powerlaw(n, n, 2.3, 2, 2000, window=1024, local_fraction=0.8,
                                            vertex_weight_min=1, vertex_weight_max=8, seed=seed * 1000 + i)
        out.append((h, int(rng.choice(KS)), 1000 + i))
    return out


def alone(ctx, host, k, seed, eps=0.03):
    dh = mtk_amd.Hypergraph(ctx, host=host)
    parts, st = dh.partition(k, eps, seed)
    dh.close()
    return parts, st


def check_balance(host, k, parts, eps=0.03):
    vw = host.arrays()["vw"]
    W = int(vw.sum())
    perfect = -(-W // k)
    bw = np.bincount(parts, weights=vw, minlength=k)
    assert bw.max() <= int((1.0 + eps) * perfect)


def test_results_identical_alone_and_in_the_mix(ctx):
    reqs = requests(12, 3)
    # host_threads: requests admitted beside few others use more initial
    # partitioning threads than alone, which must not change their result
    srv = mtk_amd.Server(0, slots=4, host_threads=16)
    tickets = [srv.submit(h, k, 0.03, sd) for h, k, sd in reqs]
    results = [srv.wait(t) for t in tickets]
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["max_in_flight"] >= 2  # requests overlapped on the device
    for (h, k, sd), (parts, res) in zip(reqs, results):
        ap, ast = alone(ctx, h, k, sd)
        assert np.array_equal(parts, ap)
        assert res["km1"] == ast["km1"]
        check_balance(h, k, parts)
        assert 0 < res["peak_bytes"] <= res["estimate_bytes"]
        assert res["latency_ms"] >= res["service_ms"] > 0
    for h, _, _ in reqs:
        h.close()


def test_tight_budget_serialises_admission(ctx):
    reqs = requests(6, 5, pmin=50_000, pmax=120_000)
    est = [mtk_amd.request_estimate(h.info().num_vertices, h.info().num_nets, h.info().num_pins,
                                    h.info().max_net_size, k) for h, k, _ in reqs]
    budget = max(est) + min(est) // 2  # never room for two requests at once
    srv = mtk_amd.Server(0, slots=4, hbm_budget_bytes=budget)
    tickets = [srv.submit(h, k, 0.03, sd) for h, k, sd in reqs]
    results = [srv.wait(t) for t in tickets]
    st = srv.stats()
    with pytest.raises(RuntimeError, match="budget"):
        big = mtk_amd.HostHypergraph.powerlaw(200_000, 200_000, 2.3, 2, 2000, seed=9)
        try:
            srv.submit(big, 8, 0.03, 1)
        finally:
            big.close()
    srv.close()
    assert st["max_in_flight"] == 1 and st["max_reserved_bytes"] <= budget
    for (h, k, sd), (parts, res) in zip(reqs, results):
        ap, ast = alone(ctx, h, k, sd)
        assert np.array_equal(parts, ap) and res["km1"] == ast["km1"]
    for h, _, _ in reqs:
        h.close()


def test_poisson_arrivals(ctx):
    reqs = requests(10, 7, pmin=3_000, pmax=60_000)
    rng = np.random.default_rng(1)
    gaps = rng.exponential(1.0 / 200.0, len(reqs))  # 200 requests per second offered
    srv = mtk_amd.Server(0, slots=4)
    t0 = time.perf_counter()
    tickets = []
    for (h, k, sd), gap in zip(reqs, np.cumsum(gaps)):
        delay = t0 + gap - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        tickets.append(srv.submit(h, k, 0.03, sd))
    lat = [srv.wait(t)[1]["latency_ms"] for t in tickets]
    srv.close()
    assert all(x > 0 for x in lat)
    for h, _, _ in reqs:
        h.close()
