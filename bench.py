"""Benchmark of the Louvain + km1 label propagation hot path on one MI355X.

Workload (default): seeded synthetic hypergraph with 1,000,000 vertices and
2,000,000 nets, net sizes ~ s^-2.2 on [2, 10000], 80% of the pins inside a
4096-wide vertex-id window, unit weights, k = 8, eps = 0.03.  The hypergraph
is resident in HBM before the timed region.

One step = Louvain community detection on the star expansion (all levels,
fast-preset limits) + re-initialisation of the partition state from the
resident starting partition (pin counts, connectivity, block weights) + km1
label propagation (5 rounds).  value = (Louvain pins swept + LP incidences
visited) / wall time of the timed steps.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "mt-kahypar_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import mtk_amd  # noqa: E402

HBM_PEAK_GBS = 8000.0
COPY_BYTES = 1 << 30  # calibration / ceiling copy: 1 GiB read + 1 GiB written
PROFILES = os.path.join(ROOT, "profiles")

WORKLOADS = {
    "powerlaw_1M_2M_k8": dict(num_vertices=1_000_000, num_nets=2_000_000, exponent=2.2, min_size=2,
                              max_size=10_000, window=4096, local_fraction=0.8, k=8, eps=0.03, seed=42),
    "powerlaw_16k_16k_k2": dict(num_vertices=16_384, num_nets=16_384, exponent=2.5, min_size=2, max_size=64,
                                window=1024, local_fraction=0.8, k=2, eps=0.03, seed=42),
    # row-net model: rows are vertices (weight = row nnz), columns are nets
    "rownet_4M_4M_k64": dict(rows=4_000_000, cols=4_000_000, mean_nnz=8.0, max_nnz=5000, band=10_000,
                             band_fraction=0.7, k=64, eps=0.03, seed=42),
}


# batched form: many independent hypergraphs refined in the same launches;
# member vertex counts log-uniform in [n_min, n_max], nets = net_ratio * n
BATCH_WORKLOADS = {
    "batch_4096_k2-16": dict(members=4096, n_min=1000, n_max=64000, net_ratio=1.2, exponent=2.5, min_size=2,
                             max_size=256, window=512, local_fraction=0.8, ks=(2, 4, 8, 16), eps=0.03, seed=42,
                             loop_stride=32, cpu_stride=8),
}


# served form: partition requests arriving as a Poisson stream; pins per
# request log-uniform in [p_min, p_max], as many vertices as nets
SERVE_WORKLOADS = {
    "serve_poisson_10k-20M": dict(requests=160, p_min=10_000, p_max=20_000_000, exponent=2.3, min_size=2,
                                  max_size=10_000, window=4096, local_fraction=0.8, vertex_weight_max=8,
                                  ks=(2, 8, 32, 128), eps=0.03, seed=42, slots=16, hw_queues=16,
                                  loads=(0.5, 1.0, 2.0), plain_stride=4),
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(name):
    w = WORKLOADS[name]
    if "rows" in w:
        return w, mtk_amd.HostHypergraph.rownet(w["rows"], w["cols"], w["mean_nnz"], w["max_nnz"], w["band"],
                                                w["band_fraction"], seed=w["seed"])
    h = mtk_amd.HostHypergraph.powerlaw(w["num_vertices"], w["num_nets"], w["exponent"], w["min_size"],
                                        w["max_size"], window=w["window"], local_fraction=w["local_fraction"],
                                        seed=w["seed"])
    return w, h


def apply_tuning(ctx, args):
    for kv in args.tuning:
        key, value = kv.split("=", 1)
        ctx.set_tuning(key, int(value))


def start_partition(n, k):
    # contiguous vertex-id blocks: with id-local pins this is a projected-like
    # starting partition whose block borders carry the LP work
    return (np.arange(n, dtype=np.int64) * k // n).astype(np.int32)


def host_info():
    """Host cores this process may use and the CPU model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    # the worker pool follows the process's CPU share (OMP_NUM_THREADS where
    # the environment sets it), never more than the cores it may run on
    share = int(os.environ.get("OMP_NUM_THREADS", affinity) or affinity)
    return dict(nproc=os.cpu_count(), affinity=affinity, threads=max(1, min(affinity, share)), cpu_model=model)


def cpu_step(oh, g, w, part, maxw, perfect, seed, threads):
    """One step of the oracle's reference-semantics path (asynchronous Louvain
    and LP, OpenMP) on the same hypergraph and starting partition."""
    k = w["k"]
    t0 = time.perf_counter()
    _, lst = g.louvain(sync=False, seed=seed, threads=threads)
    t1 = time.perf_counter()
    _, km1, st = oh.lp_refine(k, part, [maxw] * k, [perfect] * k, seed=seed, sync=False, threads=threads)
    t2 = time.perf_counter()
    return dict(louvain_s=t1 - t0, lp_s=t2 - t1, pins=lst["arcs"] // 2 + st["visits"], km1=km1)


def cpu_baseline(w, arrays, part, maxw, perfect, threads, seeds=(42, 43, 44, 45, 46)):
    """The step of the oracle's reference-semantics path over partitioner
    seeds 42-46 on `threads` host threads (median reported), and seed 42 on
    one thread for context."""
    import oracle as orc
    oh = orc.Hypergraph(arrays["n"], arrays["noff"], arrays["pins"], arrays["vw"], arrays["nw"])
    g = orc.Graph(oh, orc.Graph.UNIFORM, ranked=True)
    runs = [cpu_step(oh, g, w, part, maxw, perfect, sd, threads) for sd in seeds]
    one = cpu_step(oh, g, w, part, maxw, perfect, seeds[0], 1)
    rates = [r["pins"] / (r["louvain_s"] + r["lp_s"]) for r in runs]
    med = sorted(range(len(runs)), key=lambda i: rates[i])[len(runs) // 2]
    return dict(runs=runs, median=runs[med], median_rate=rates[med], one_thread=one,
                one_thread_rate=one["pins"] / (one["louvain_s"] + one["lp_s"]))


def lp_parity(arrays, part, k, maxw, perfect, seed, gpu_parts, gpu_km1):
    """The device LP's block ids and km1 against the oracle's synchronous mode
    from the same starting partition (bit for bit)."""
    import oracle as orc
    oh = orc.Hypergraph(arrays["n"], arrays["noff"], arrays["pins"], arrays["vw"], arrays["nw"])
    t0 = time.perf_counter()
    rpart, rkm1, _ = oh.lp_refine(k, part, [maxw] * k, [perfect] * k, seed=seed, sync=True)
    ok = bool(np.array_equal(rpart, gpu_parts)) and int(rkm1) == int(gpu_km1)
    return dict(lp_vs_sync_oracle="bit-exact" if ok else "MISMATCH", km1_gpu=int(gpu_km1), km1_oracle=int(rkm1),
                oracle_s=round(time.perf_counter() - t0, 2))


def phases_of(ctx, dh, w, arrays, cpu, threads):
    """One full pipeline run (Louvain, coarsening, initial partitioning, LP
    on every level) of the same hypergraph with every kernel family timed:
    device time per phase, the rating roofline, and the CPU restatement's
    time per phase (Louvain and LP: the baseline step's median; coarsening:
    the reference-semantics rater's throughput on a sample of level-0
    vertices, scaled to the pipeline's rating work; initial partitioning:
    the CPU restatement of the device partitioner on the coarsest level of a
    coarsening of the same hypergraph)."""
    import oracle as orc
    ctx.set_profiling(True)
    ctx.reset_profile()
    part, st = dh.partition(w["k"], w["eps"], w["seed"])
    fam = ctx.profile()
    kern = ctx.kernel_profile()
    ctx.set_profiling(False)
    ctx.reset_profile()
    n = arrays["n"]
    oh = orc.Hypergraph(n, arrays["noff"], arrays["pins"], arrays["vw"], arrays["nw"])
    sample = np.arange(0, n, 50, dtype=np.uint32)
    t0 = time.perf_counter()
    vis = oh.rating_sample_ref(sample, threads=threads)
    rate = vis / max(time.perf_counter() - t0, 1e-9)
    rating = fam["rating"]
    contraction = fam["contraction"]
    ip_cpu_ms, ip_sample = None, None
    if cpu:
        # the oracle's initial partitioning (same rules, host threads) on the
        # coarsest level of the device coarsening of the same hypergraph
        comm, _ = dh.louvain(seed=w["seed"])
        dh.set_communities(comm)
        W = int(arrays["vw"].sum())
        climit = 160 * w["k"]
        hi = dh.coarsen(contraction_limit=climit, max_node_weight=int(-(-W // climit)), seed=w["seed"])
        ca = hi.level(hi.num_levels() - 1).arrays()
        cn = len(ca["vw"])
        maxw = int((1.0 + w["eps"]) * -(-W // w["k"]))
        coh = orc.Hypergraph(cn, ca["noff"], ca["pins"], ca["vw"], ca["nw"])
        t0 = time.perf_counter()
        coh.initial_partition(w["k"], [maxw] * w["k"], w["seed"], threads=threads)
        ip_cpu_ms = round(1e3 * (time.perf_counter() - t0), 1)
        ip_sample = "oracle initial partitioning (20 runs per bisection, %d threads) of a %d-vertex, %d-net " \
                    "coarsest level" % (threads, cn, len(ca["nw"]))
        hi.close()
    out = {
        "workload_note": "one multilevel pipeline run of the same hypergraph (k=%d, eps=%g, seed %d): km1 %d, "
                         "imbalance %.4f, %d levels" % (w["k"], w["eps"], w["seed"], st["km1"], st["imbalance"],
                                                       st["levels"]),
        "louvain": {"gpu_ms": round(st["louvain_ms"], 2),
                    "cpu_ms": round(1e3 * cpu["median"]["louvain_s"], 1) if cpu else None},
        "coarsening": {"gpu_ms": round(st["coarsening_ms"], 2),
                       "rating_gpu_ms": round(rating["ms"], 2), "contraction_gpu_ms": round(contraction["ms"], 2),
                       "rating_visits": st["rating_visits"],
                       "rating_roofline": {"bound": "hbm", "achieved": round(rating["bytes"] / (rating["ms"] * 1e6), 1)
                                           if rating["ms"] else None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                           "frac": round(rating["bytes"] / (rating["ms"] * 1e6) / HBM_PEAK_GBS, 4)
                                           if rating["ms"] else None,
                                           "bytes": "8 per rating pin visit + 12 per pin + 12 per vertex, per pass"},
                       "cpu_ms_estimate": round(1e3 * st["rating_visits"] / rate, 1),
                       "cpu_sample": "reference-semantics rating of every 50th level-0 vertex on %d threads: %d pin "
                                     "visits at %.3g visits/s, scaled to the pipeline's %d rating visits" % (
                                         threads, vis, rate, st["rating_visits"])},
        "initial_partitioning": {"gpu_ms": round(st["initial_ms"], 2),
                                 "kernel_ms": {k: round(v["ms"], 2) for k, v in kern.items()
                                               if k.startswith("ip_") and v["launches"]},
                                 "cpu_ms": ip_cpu_ms, "cpu_sample": ip_sample},
        "lp": {"gpu_ms_all_levels": round(st["refinement_ms"], 2),
               "cpu_ms_finest_level": round(1e3 * cpu["median"]["lp_s"], 1) if cpu else None},
        "total_gpu_pipeline_ms": round(st["total_ms"], 2),
    }
    # FM (10 multitry rounds, 25 seeds per search) on the top level of the
    # pipeline's LP-refined partition: km1 with LP + FM
    maxw = int((1.0 + w["eps"]) * -(-int(arrays["vw"].sum()) // w["k"]))
    dp = mtk_amd.Partition(dh, w["k"])
    dp.set_parts(part)
    km1_lp = dp.km1()
    ctx.set_profiling(True)
    ctx.reset_profile()
    fst = dp.fm_refine([maxw] * w["k"], seed=w["seed"], rounds=10)
    kp = ctx.kernel_profile()
    ctx.set_profiling(False)
    ctx.reset_profile()
    km1_fm = dp.km1()
    pw = np.bincount(dp.parts(), weights=arrays["vw"], minlength=w["k"])
    dp.close()
    out["fm_top_level"] = {
        "km1_lp": km1_lp, "km1_lp_fm": km1_fm, "improvement": fst["improvement"],
        "balanced": bool(pw.max() <= maxw), "wall_ms": round(fst["device_ms"], 1),
        "rounds": fst["rounds"], "searches": fst["searches"], "moves": fst["moves"],
        "kernel_ms": {name: round(v["ms"], 2) for name, v in kp.items() if name.startswith("fm_")},
        "note": "FM is off in the default pipeline (fm_rounds = 0) and on with r-fm-type=fm_gain_cache; "
                "every level with FM: fm_rounds = 10",
    }
    return out


def counter_summary(workload, kernel):
    """From the committed counter summary of the workload: HBM bytes per
    launch of `kernel` (rocprofv3 FETCH_SIZE and WRITE_SIZE runs, each scaled
    by the factor the same counter shows for the calibration copy) and its
    device atomics per launch (TCC_ATOMIC_sum), each None when absent."""
    path = os.path.join(PROFILES, "hbm_traffic_%s.json" % workload)
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        t = json.load(f)
    k = t.get("kernels", {}).get(kernel)
    if not k:
        return None, None
    traffic = round(k["fetch_bytes"] / t["fetch_factor"] + k["write_bytes"] / t["write_factor"])
    return traffic, k.get("atomics")


def normalize_kernel_name(name):
    """'void mkp::f<mkp::StarView, 32>(args)' -> 'f<StarView, 32>'."""
    name = name.strip().strip('"')
    if name.startswith("void "):
        name = name[5:]
    depth, cut = 0, len(name)
    for i, ch in enumerate(name):
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            cut = i
            break
    return name[:cut].replace("mkp::", "").strip()


def pmc_summary(fetch_dir, write_dir, workload, atomic_dir=None):
    """Per-kernel mean FETCH_SIZE / WRITE_SIZE per dispatch (bytes) from two
    rocprofv3 --pmc runs of this script, the factors the calibration copy
    (COPY_BYTES read and written) gives and, from a third run, TCC_ATOMIC_sum
    per dispatch; written to profiles/."""
    import csv
    import glob

    def load(d, counter):
        acc = {}
        for fn in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(fn) as f:
                for row in csv.DictReader(f):
                    if row.get("Counter_Name") != counter:
                        continue
                    k = normalize_kernel_name(row["Kernel_Name"])
                    a = acc.setdefault(k, [0.0, 0])
                    a[0] += float(row["Counter_Value"]) * 1024.0  # KiB per dispatch
                    a[1] += 1
        return {k: v[0] / v[1] for k, v in acc.items() if v[1]}

    fetch, write = load(fetch_dir, "FETCH_SIZE"), load(write_dir, "WRITE_SIZE")
    atom = {k: v / 1024.0 for k, v in load(atomic_dir, "TCC_ATOMIC_sum").items()} if atomic_dir else {}
    copy = "stream_copy_kernel"
    out = dict(workload=workload, fetch_factor=fetch[copy] / COPY_BYTES, write_factor=write[copy] / COPY_BYTES,
               calibration="stream_copy_kernel: %d bytes read and %d written per dispatch" % (COPY_BYTES, COPY_BYTES),
               kernels={k: dict(fetch_bytes=fetch.get(k, 0.0), write_bytes=write.get(k, 0.0),
                                **({"atomics": atom[k]} if k in atom else {}))
                        for k in sorted(set(fetch) | set(write))})
    path = os.path.join(PROFILES, "hbm_traffic_%s.json" % workload)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


def dominant_kernel(ctx):
    """The single kernel (template instantiation) with the most device time
    among those with an algorithmic byte count, from the per-launch event
    timings of every kernel (logged as a table)."""
    kern = {k: v for k, v in ctx.kernel_profile().items() if v["launches"] and v["bytes"] > 0}
    for name, v in sorted(kern.items(), key=lambda kv: -kv[1]["ms"]):
        log("  kernel %-52s %8.3f ms %5d launches %8.1f us/launch %9.1f GB/s" % (
            name, v["ms"], v["launches"], 1e3 * v["ms"] / v["launches"], v["bytes"] / (v["ms"] * 1e6)))
    return max(kern, key=lambda k: kern[k]["ms"])


def roofline_of(ctx, prof, workload, dom=None):
    """Roofline entry of the dominant kernel; every launch of it was timed on
    its own with an event pair on the launch stream.  With `dom` given, the
    profile holds that kernel's launches of the timed steps only."""
    kern = {k: v for k, v in ctx.kernel_profile().items() if v["launches"] and v["bytes"] > 0}
    if dom is None:
        dom = dominant_kernel(ctx)
    d = kern[dom]
    avg_ms = d["ms"] / d["launches"]
    bytes_per_launch = d["bytes"] / d["launches"]
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic, atomics = counter_summary(workload, dom)
    copy_ms = ctx.stream_copy(COPY_BYTES, 20)
    return dict(bound="hbm", kernel=dom, achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                atomics_per_launch=round(atomics) if atomics is not None else None,
                atomic_rate_per_s=round(atomics / (avg_ms * 1e-3), 1) if atomics is not None else None,
                avg_launch_us=round(avg_ms * 1e3, 2), launches=d["launches"],
                algorithmic_bytes_per_launch=round(bytes_per_launch),
                share_of_step_device_time=(round(d["ms"] / sum(v["ms"] for v in prof.values()), 3)
                                           if prof else None),
                measured_copy_GBs=round(2 * COPY_BYTES / (copy_ms * 1e-3) / 1e9, 1))


def make_batch(w):
    """Member hypergraphs of a batched workload (generated on host threads)."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(w["seed"])
    G = w["members"]
    ns = np.exp(rng.uniform(np.log(w["n_min"]), np.log(w["n_max"]), G)).astype(np.int64)
    ks = [int(k) for k in rng.choice(w["ks"], G)]

    def gen(g):
        n = int(ns[g])
        return mtk_amd.HostHypergraph.powerlaw(n, int(round(w["net_ratio"] * n)), w["exponent"], w["min_size"],
                                               w["max_size"], window=w["window"], local_fraction=w["local_fraction"],
                                               seed=w["seed"] + g)

    mtk_amd.lib()  # load the library before the worker threads use it
    with ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1))) as ex:
        hosts = list(ex.map(gen, range(G)))
    return hosts, ks


def batch_cpu_baseline(w, hosts, ks, sample, threads):
    """The oracle's reference-semantics path (asynchronous Louvain and LP,
    OpenMP) on a sample of the members, one after another."""
    import oracle as orc
    pins, secs = 0, 0.0
    for g in sample:
        a = hosts[g].arrays()
        n, k = a["n"], ks[g]
        W = int(a["vw"].sum())
        perfect = -(-W // k)
        maxw = int((1.0 + w["eps"]) * perfect)
        oh = orc.Hypergraph(n, a["noff"], a["pins"], a["vw"], a["nw"])
        gr = orc.Graph(oh, orc.Graph.UNIFORM, ranked=True)
        t0 = time.perf_counter()
        _, lst = gr.louvain(sync=False, seed=w["seed"] + g, threads=threads)
        _, _, st = oh.lp_refine(k, start_partition(n, k), [maxw] * k, [perfect] * k, seed=w["seed"] + g,
                                sync=False, threads=threads)
        secs += time.perf_counter() - t0
        pins += lst["arcs"] // 2 + st["visits"]
    return pins, secs


def batch_main(args):
    """One step = Louvain of every member (all levels) + re-initialisation of
    every member's partition state from its resident starting partition + km1
    LP of every member, all members in the same launches.  Beside it, the
    plain one-hypergraph-per-call GPU path over every loop_stride-th member."""
    w = BATCH_WORKLOADS[args.workload]
    t_gen = time.perf_counter()
    hosts, ks = make_batch(w)
    infos = [h.info() for h in hosts]
    ns = [i.num_vertices for i in infos]
    G = len(hosts)
    P = sum(i.num_pins for i in infos)
    log("workload %s: %d members, n=%d m=%d pins=%d (generated in %.1f s)" % (
        args.workload, G, sum(ns), sum(i.num_nets for i in infos), P, time.perf_counter() - t_gen))
    ctx = mtk_amd.Context(0)
    apply_tuning(ctx, args)
    b = mtk_amd.Batch(ctx, hosts)
    seeds = [w["seed"] + g for g in range(G)]
    part0 = np.concatenate([start_partition(n, k) for n, k in zip(ns, ks)])
    base = mtk_amd.BatchPartition(b, ks)
    base.set_parts(part0)
    km1_start = base.km1()
    work = mtk_amd.BatchPartition(b, ks)

    def step():
        _, lst = b.louvain(seeds=seeds, want_communities=False)
        work.copy_parts_from(base)
        st = work.lp_refine(seeds, epsilon=w["eps"])
        return sum(x["arcs_swept"] for x in lst) // 2 + sum(x["visits"] for x in st)

    ctx.set_profiling(True)
    ctx.reset_profile()
    for i in range(args.warmup):
        step()
        log("warmup step %d done" % i)
    ctx.synchronize()
    dom = share = None
    if args.warmup:  # as in the single-hypergraph bench: only the dominant kernel is timed below
        prof = ctx.profile()
        for f, v in prof.items():
            if v["launches"]:
                log("  %-18s %9.3f ms %6d launches %10.3f GB/s (warm-up)" % (
                    f, v["ms"], v["launches"], v["bytes"] / (v["ms"] * 1e6) if v["ms"] else 0))
        dom = dominant_kernel(ctx)
        share = ctx.kernel_profile()[dom]["ms"] / sum(v["ms"] for v in prof.values())
        ctx.reset_profile()
        ctx.set_profile_filter(dom)
    units = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        units += step()
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.set_profile_filter(None)
    km1_end = work.km1()
    log("timed %d steps in %.3f s; km1 (sum over members) %d -> %d" % (args.steps, wall, km1_start.sum(),
                                                                      km1_end.sum()))
    if dom is None:
        roofline = roofline_of(ctx, ctx.profile(), args.workload)
    else:
        roofline = roofline_of(ctx, None, args.workload, dom)
        roofline["share_of_step_device_time"] = round(share, 3)
    ctx.set_profiling(False)

    # plain path: one hypergraph per call, same members, seeds and settings
    sample = [] if args.no_loop else list(range(0, G, w["loop_stride"]))
    dev = []
    for g in sample:
        n, k = ns[g], ks[g]
        dh = mtk_amd.Hypergraph(ctx, host=hosts[g])
        perfect = -(-int(infos[g].total_weight) // k)
        maxw = int((1.0 + w["eps"]) * perfect)
        bp = mtk_amd.Partition(dh, k)
        bp.set_parts(start_partition(n, k))
        dev.append((dh, bp, mtk_amd.Partition(dh, k), k, maxw, perfect, seeds[g]))

    def loop_step():
        u = 0
        for dh, bp, wp, k, maxw, perfect, sd in dev:
            _, lst = dh.louvain(seed=sd, want_communities=False)
            wp.copy_parts_from(bp)
            st = wp.lp_refine([maxw] * k, [perfect] * k, seed=sd)
            u += lst["arcs_swept"] // 2 + st["visits"]
        return u

    loop_wall = loop_units = same = None
    if sample:
        loop_step()
        ctx.synchronize()
        t1 = time.perf_counter()
        loop_units = loop_step()
        ctx.synchronize()
        loop_wall = time.perf_counter() - t1
        same = all(int(d[2].km1()) == int(km1_end[g]) for d, g in zip(dev, sample))
        log("per-hypergraph loop over %d members: %.3f s, km1 equal to the batch: %s" % (len(sample), loop_wall,
                                                                                         same))
    for d in dev:
        d[2].close()
        d[1].close()
        d[0].close()

    cpu = None
    if not args.no_cpu:
        threads = host_info()["threads"]
        csample = list(range(0, G, w["cpu_stride"]))
        log("cpu baseline: oracle reference-semantics Louvain + LP on %d members, %d threads" % (len(csample),
                                                                                               threads))
        cp, cs = batch_cpu_baseline(w, hosts, ks, csample, threads)
        cpu = dict(value=round(cp / cs, 1), unit="pins/s", cores=threads, kind="port",
                   sample="every %d-th member (%d members), one after another: %.2f s (asynchronous restatement, "
                          "OpenMP)" % (w["cpu_stride"], len(csample), cs))

    hg_per_s = G * args.steps / wall
    loop_hg_per_s = len(sample) / loop_wall if sample else None
    out = {
        "metric": "LP+Louvain pins/s",
        "value": round(units / wall, 1),
        "unit": "pins/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": args.workload, "members": G, "vertices": sum(ns),
                   "nets": sum(i.num_nets for i in infos), "pins": P, "k": "per member from %s" % (w["ks"],),
                   "epsilon": w["eps"], "km1_start_sum": int(km1_start.sum()), "km1_end_sum": int(km1_end.sum()),
                   "parallelism": "single GPU"},
        "batched": {"hypergraphs_per_s": round(hg_per_s, 1),
                    "loop_hypergraphs_per_s": round(loop_hg_per_s, 1) if sample else None,
                    "loop_pins_per_s": round(loop_units / loop_wall, 1) if sample else None,
                    "speedup_vs_loop": round(hg_per_s / loop_hg_per_s, 2) if sample else None,
                    "loop_sample": "every %d-th member (%d members)" % (w["loop_stride"], len(sample)),
                    "km1_equal_to_loop": same},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    work.close()
    base.close()
    b.close()


def powerlaw_mean(exponent, lo, hi):
    s = np.arange(lo, hi + 1, dtype=np.float64)
    w = s ** -exponent
    return float((s * w).sum() / w.sum())


def make_requests(w):
    """Request hypergraphs of a served workload (generated on host threads)."""
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(w["seed"])
    R = w["requests"]
    ps = np.exp(rng.uniform(np.log(w["p_min"]), np.log(w["p_max"]), R))
    ks = [int(k) for k in rng.choice(w["ks"], R)]
    mean = powerlaw_mean(w["exponent"], w["min_size"], w["max_size"])

    def gen(i):
        n = max(64, int(round(ps[i] / mean)))
        return mtk_amd.HostHypergraph.powerlaw(n, n, w["exponent"], w["min_size"], w["max_size"], window=w["window"],
                                               local_fraction=w["local_fraction"], vertex_weight_min=1,
                                               vertex_weight_max=w["vertex_weight_max"], seed=w["seed"] + i)

    mtk_amd.lib()
    with ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1))) as ex:
        hosts = list(ex.map(gen, range(R)))
    return hosts, ks


def serve_run(w, hosts, ks, arrivals):
    """Submits request i at arrivals[i] seconds (None: all at once) and returns
    (makespan s, per-request results, server stats)."""
    srv = mtk_amd.Server(0, slots=w["slots"])
    order = list(range(len(hosts)))
    t0 = time.perf_counter()
    tickets = []
    for i in order:
        if arrivals is not None:
            d = t0 + arrivals[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
        tickets.append(srv.submit(hosts[i], ks[i], w["eps"], w["seed"] + i))
    res = []
    for j, t in enumerate(tickets):
        res.append(srv.wait(t, want_parts=False)[1])
        if (j + 1) % 40 == 0:
            log("  %d of %d requests done (%.1f s)" % (j + 1, len(tickets), time.perf_counter() - t0))
    makespan = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return makespan, res, st


def serve_main(args):
    """Served form: saturation throughput (every request submitted at once),
    then Poisson arrivals at fractions of it with p50 / p99 latency, beside the
    plain path (one request at a time on one stream, upload included) over
    every plain_stride-th request."""
    w = SERVE_WORKLOADS[args.workload]
    t_gen = time.perf_counter()
    hosts, ks = make_requests(w)
    infos = [h.info() for h in hosts]
    R = len(hosts)
    P = sum(i.num_pins for i in infos)
    log("workload %s: %d requests, %d pins in total (mean %.0f), generated in %.1f s" % (
        args.workload, R, P, P / R, time.perf_counter() - t_gen))

    # plain path
    ctx = mtk_amd.Context(0)
    apply_tuning(ctx, args)
    sample = list(range(0, R, w["plain_stride"]))
    plain_km1 = {}
    for i in sample[:2]:  # warm-up of the plain path
        dh = mtk_amd.Hypergraph(ctx, host=hosts[i])
        dh.partition(ks[i], w["eps"], w["seed"] + i)
        dh.close()
    ctx.synchronize()
    t0 = time.perf_counter()
    for j, i in enumerate(sample):
        dh = mtk_amd.Hypergraph(ctx, host=hosts[i])
        _, st = dh.partition(ks[i], w["eps"], w["seed"] + i)
        plain_km1[i] = st["km1"]
        dh.close()
        if (j + 1) % 10 == 0:
            log("  plain path: %d of %d requests (%.1f s)" % (j + 1, len(sample), time.perf_counter() - t0))
    ctx.synchronize()
    plain_s = time.perf_counter() - t0
    plain_rate = len(sample) / plain_s
    plain_pins = sum(infos[i].num_pins for i in sample) / plain_s
    ctx.close()
    log("plain path: %d requests in %.2f s = %.2f req/s" % (len(sample), plain_s, plain_rate))

    serve_run(w, hosts[:2 * w["slots"]], ks[:2 * w["slots"]], None)  # warm-up of the slots
    makespan, res, st = serve_run(w, hosts, ks, None)
    sat = R / makespan
    same = all(res[i]["km1"] == plain_km1[i] for i in sample)
    log("saturation: %d requests in %.2f s = %.2f req/s (max in flight %d); km1 equal to the plain path: %s" % (
        R, makespan, sat, st["max_in_flight"], same))
    peak_over = sum(1 for r in res if r["peak_bytes"] > r["estimate_bytes"])

    loads = []
    rng = np.random.default_rng(w["seed"] + 1)
    for f in w["loads"]:
        qps = f * sat
        arrivals = np.cumsum(rng.exponential(1.0 / qps, R))
        mk, rr, _ = serve_run(w, hosts, ks, arrivals)
        lat = np.array([r["latency_ms"] for r in rr])
        loads.append(dict(offered_fraction_of_saturation=f, offered_req_per_s=round(qps, 2),
                          achieved_req_per_s=round(R / mk, 2), p50_ms=round(float(np.percentile(lat, 50)), 2),
                          p99_ms=round(float(np.percentile(lat, 99)), 2),
                          mean_queue_ms=round(float(np.mean([r["queue_ms"] for r in rr])), 2)))
        log("load %.1fx: offered %.2f req/s, achieved %.2f, p50 %.1f ms, p99 %.1f ms" % (
            f, qps, R / mk, loads[-1]["p50_ms"], loads[-1]["p99_ms"]))

    out = {
        "metric": "served req/s",
        "value": round(sat, 2),
        "unit": "req/s",
        "n_gpus": 1,
        "steps": R,
        "warmup": 2 * w["slots"],
        "ms_per_step": round(makespan / R * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": args.workload, "requests": R, "pins_total": P, "pins_per_request":
                   "log-uniform [%d, %d]" % (w["p_min"], w["p_max"]), "k": "per request from %s" % (w["ks"],),
                   "epsilon": w["eps"], "slots": w["slots"],
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")), "parallelism": "single GPU"},
        "served": {"saturation_req_per_s": round(sat, 2), "saturation_pins_per_s": round(P / makespan, 1),
                   "plain_req_per_s": round(plain_rate, 2), "plain_pins_per_s": round(plain_pins, 1),
                   "speedup_vs_plain": round(sat / plain_rate, 2),
                   "plain_sample": "every %d-th request (%d requests)" % (w["plain_stride"], len(sample)),
                   "km1_equal_to_plain": same, "max_in_flight": st["max_in_flight"],
                   "max_reserved_bytes": st["max_reserved_bytes"], "budget_bytes": st["budget_bytes"],
                   "requests_above_estimate": peak_over, "loads": loads},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    for h in hosts:
        h.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="powerlaw_1M_2M_k8", choices=sorted(WORKLOADS) + sorted(BATCH_WORKLOADS) + sorted(SERVE_WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-loop", action="store_true",
                    help="batched workloads: skip the per-hypergraph comparison leg (for per-kernel profiles)")
    ap.add_argument("--pmc-summary", nargs="+", metavar="DIR",
                    help="summarise rocprofv3 --pmc output folders of this script (FETCH_SIZE, WRITE_SIZE and "
                         "optionally TCC_ATOMIC_sum runs) into profiles/ and exit")
    ap.add_argument("--no-phases", action="store_true", help="skip the per-phase pipeline run")
    ap.add_argument("--tuning", action="append", default=[], metavar="KEY=VALUE",
                    help="kernel shape choices of the context (mkp_context_set_tuning), e.g. "
                         "louvain_fused_classes=0 gives every degree class of a Louvain sub-round its own launch")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(args.pmc_summary[0], args.pmc_summary[1], args.workload,
                    args.pmc_summary[2] if len(args.pmc_summary) > 2 else None)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (requested --gpus %d)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload in BATCH_WORKLOADS:
        batch_main(args)
        return
    if args.workload in SERVE_WORKLOADS:
        # every slot owns a stream; HIP maps the streams of a process onto
        # GPU_MAX_HW_QUEUES hardware queues (4 by default), and streams that
        # share a queue run one after the other.  Raised to the workload's
        # count before the first HIP call of the process (the runtime reads
        # it when it starts).
        want = SERVE_WORKLOADS[args.workload]["hw_queues"]
        try:
            have = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
        except ValueError:
            have = 4
        if have < want:
            os.environ["GPU_MAX_HW_QUEUES"] = str(want)
        serve_main(args)
        return

    w, host = make_workload(args.workload)
    arrays = host.arrays()
    n, k = arrays["n"], w["k"]
    P = len(arrays["pins"])
    log("workload %s: n=%d m=%d pins=%d" % (args.workload, n, len(arrays["noff"]) - 1, P))
    ctx = mtk_amd.Context(0)
    apply_tuning(ctx, args)
    dh = mtk_amd.Hypergraph(ctx, host=host)
    W = int(arrays["vw"].sum())
    perfect = -(-W // k)
    maxw = int((1.0 + w["eps"]) * perfect)
    part0 = start_partition(n, k)
    base = mtk_amd.Partition(dh, k)
    base.set_parts(part0)
    km1_start = base.km1()
    work = mtk_amd.Partition(dh, k)

    def step():
        _, lst = dh.louvain(seed=w["seed"], want_communities=False)
        work.copy_parts_from(base)
        st = work.lp_refine([maxw] * k, [perfect] * k, seed=w["seed"])
        # Louvain work in pins: two arcs of the star expansion per pin
        return lst["arcs_swept"] // 2 + st["visits"], (lst, st)

    # warm-up steps with every kernel timed: the per-family table and the
    # dominant kernel; the timed steps then carry event pairs around the
    # launches of the dominant kernel only
    # (without warm-up steps every kernel is timed inside the timed steps)
    ctx.set_profiling(True)
    ctx.reset_profile()
    for i in range(args.warmup):
        step()
        log("warmup step %d done" % i)
    ctx.synchronize()
    dom = share = None
    if args.warmup:
        prof = ctx.profile()
        for f, v in prof.items():
            if v["launches"]:
                log("  %-18s %9.3f ms %6d launches %10.3f GB/s (warm-up)" % (
                    f, v["ms"], v["launches"], v["bytes"] / (v["ms"] * 1e6) if v["ms"] else 0))
        dom = dominant_kernel(ctx)
        share = ctx.kernel_profile()[dom]["ms"] / sum(v["ms"] for v in prof.values())
        ctx.reset_profile()
        ctx.set_profile_filter(dom)
    work_units = 0
    last = None
    t0 = time.perf_counter()
    for i in range(args.steps):
        u, last = step()
        work_units += u
    ctx.synchronize()
    wall = time.perf_counter() - t0
    ctx.set_profile_filter(None)
    km1_end = work.km1()
    log("timed %d steps in %.3f s; km1 %d -> %d" % (args.steps, wall, km1_start, km1_end))
    if dom is None:
        prof = ctx.profile()
        roofline = roofline_of(ctx, prof, args.workload)
    else:
        roofline = roofline_of(ctx, None, args.workload, dom)
        roofline["share_of_step_device_time"] = round(share, 3)
    ctx.set_profiling(False)
    hi = host_info()
    threads = hi["threads"]

    # the LP of the last timed step against the oracle's synchronous mode
    parity = lp_parity(arrays, part0, k, maxw, perfect, w["seed"], work.parts(), km1_end)
    log("parity: LP %s vs the synchronous oracle (km1 %d / %d)" % (parity["lp_vs_sync_oracle"], parity["km1_gpu"],
                                                                   parity["km1_oracle"]))

    cpu = cpu_run = None
    if not args.no_cpu:
        log("cpu baseline: oracle reference-semantics Louvain + LP, %d threads, seeds 42-46, and one thread" % threads)
        cpu_run = cpu_baseline(w, arrays, part0, maxw, perfect, threads)
        m = cpu_run["median"]
        cpu = dict(value=round(cpu_run["median_rate"], 1), unit="pins/s", cores=threads, kind="port",
                   sample="one step (Louvain + LP, asynchronous restatement, OpenMP) on the same hypergraph and "
                          "starting partition per partitioner seed 42-46, median reported: Louvain %.2f s + LP %.2f s, "
                          "LP km1 %d; seed 42 on one thread: %.3g pins/s (%.2f s)" % (
                              m["louvain_s"], m["lp_s"], m["km1"], cpu_run["one_thread_rate"],
                              cpu_run["one_thread"]["louvain_s"] + cpu_run["one_thread"]["lp_s"]),
                   seeds_pins_per_s=[round(r["pins"] / (r["louvain_s"] + r["lp_s"]), 1) for r in cpu_run["runs"]],
                   one_thread_pins_per_s=round(cpu_run["one_thread_rate"], 1),
                   host={"nproc": hi["nproc"], "affinity_cores": hi["affinity"], "cpu_model": hi["cpu_model"]})

    phases = None
    if not args.no_phases:
        log("phases: one pipeline run with every kernel family timed")
        phases = phases_of(ctx, dh, w, arrays, cpu_run, threads)

    value = work_units / wall
    out = {
        "metric": "LP+Louvain pins/s",
        "value": round(value, 1),
        "unit": "pins/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": args.workload, "vertices": n, "nets": len(arrays["noff"]) - 1, "pins": P,
                   "k": k, "epsilon": w["eps"], "km1_start": km1_start, "km1_end": km1_end,
                   "parallelism": "single GPU",
                   "outside_timed_region": "hypergraph upload; net locality ranks (built once per hypergraph at "
                                           "its first Louvain call, in the warm-up)"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "parity": parity,
        "phases": phases,
    }
    print(json.dumps(out), flush=True)
    if parity["lp_vs_sync_oracle"] != "bit-exact":
        sys.exit(3)


if __name__ == "__main__":
    main()
