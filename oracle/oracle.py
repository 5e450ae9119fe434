"""ctypes bindings of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  The library is built by
``make -C oracle`` into ``oracle/_build/liboracle.so``.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
vp = C.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_hg_new": (vp, [C.c_uint32, C.c_uint32, u64p, u32p, vp, vp]),
            "orc_hg_free": (None, [vp]),
            "orc_hg_set_comm": (None, [vp, i32p]),
            "orc_hg_sizes": (None, [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint32), C.POINTER(C.c_int64)]),
            "orc_hg_get": (None, [vp, u64p, u32p, u64p, u32p, i32p, i32p, i32p]),
            "orc_phg_init": (C.c_int64, [vp, C.c_int, i32p, i64p, u32p, u64p]),
            "orc_cut": (C.c_int64, [vp, C.c_int, i32p]),
            "orc_border": (None, [vp, C.c_int, i32p, u8p]),
            "orc_max_gain_move": (None, [vp, C.c_int, i32p, C.c_uint32, i64p, C.POINTER(C.c_int32),
                                         C.POINTER(C.c_int32)]),
            "orc_lp_refine": (C.c_int64, [vp, C.c_int, i32p, i64p, i64p, C.c_uint64, C.c_int, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_int, u32p, C.c_uint32, i64p]),
            "orc_fm_refine": (C.c_int64, [vp, C.c_int, i32p, i64p, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_uint32, C.c_uint64, C.c_uint64, C.c_int, i64p]),
            "orc_fm_recalculate": (None, [vp, C.c_int, i32p, C.c_uint32, u32p, i32p, i32p, i64p]),
            "orc_fm_best_prefix": (C.c_uint64, [vp, C.c_int, i32p, C.c_uint32, u32p, i32p, i32p, i64p, i64p,
                                                C.POINTER(C.c_int64)]),
            "orc_graph_new": (vp, [vp, C.c_int, C.c_int]),
            "orc_graph_free": (None, [vp]),
            "orc_graph_sizes": (None, [vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_double),
                                       C.POINTER(C.c_uint64)]),
            "orc_graph_get": (None, [vp, u64p, u32p, f64p, f64p]),
            "orc_graph_contract": (vp, [vp, i32p]),
            "orc_modularity": (C.c_double, [vp, i32p]),
            "orc_louvain_best_cluster": (C.c_int32, [vp, i32p, C.c_uint32, C.c_uint64, C.POINTER(C.c_double)]),
            "orc_louvain": (None, [vp, C.c_int, C.c_uint64, C.c_int, C.c_double, C.c_uint64, C.c_int, C.c_int,
                                   C.c_int, i32p, i64p]),
            "orc_louvain_level_sync": (C.c_int, [vp, C.c_uint64, C.c_int, C.c_double, C.c_uint64, C.c_int, i32p,
                                                 i64p]),
            "orc_cluster_pass": (C.c_uint32, [vp, C.c_int, C.c_int, C.c_uint32, C.c_int32, C.c_double, C.c_uint32,
                                              C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int, u32p, i64p]),
            "orc_hg_contract": (vp, [vp, u32p, u32p]),
            "orc_rating_sample_ref": (C.c_int64, [vp, u32p, C.c_uint32, C.c_uint32, C.c_int]),
            "orc_rate_sync": (None, [vp, C.c_int, C.c_int, C.c_int32, C.c_uint32, C.c_uint64, C.c_uint64, C.c_int,
                                     u32p, i64p, u8p, u32p, C.c_uint32, u32p, u64p]),
            "orc_louvain_level_sync_trace": (C.c_int, [vp, C.c_uint64, C.c_int, C.c_double, C.c_uint64, C.c_int,
                                                       i32p, u32p, C.c_uint64, u64p, u32p, i32p, C.c_uint64, u64p,
                                                       C.c_uint64]),
            "orc_coarsen": (C.c_int, [vp, C.c_int, C.c_uint32, C.c_int32, C.c_double, C.c_double, C.c_uint32,
                                      C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
            "orc_partition": (C.c_int64, [vp, C.c_int, C.c_double, C.c_uint64, C.c_int, C.c_int, C.c_int, i32p, f64p,
                                          i64p]),
            "orc_reduce": (vp, [vp, C.c_uint32, C.c_uint32, u32p, u32p]),
            "orc_rebalance": (C.c_int64, [vp, C.c_int, i32p, i64p, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
            "orc_large_net_threshold": (C.c_uint32, [C.c_uint32, C.c_double]),
            "orc_restore_degree_zero": (None, [C.c_uint32, u32p, C.c_uint32, i32p, C.c_int, i64p, i32p]),
            "orc_initial_partition": (C.c_int64, [vp, C.c_int, i64p, C.c_uint64, C.c_int, C.c_int, i32p]),
            "orc_ip_bisect_run": (None, [vp, C.c_int64, C.c_int64, C.c_int64, C.c_uint64, C.c_int, C.c_int, u8p,
                                         i64p]),
            "orc_ip_perm": (None, [C.c_uint64, C.c_uint32, u32p]),
            "orc_ip_stream": (C.c_uint64, [C.c_uint64, C.c_int, C.c_int]),
            "orc_ip_run_key": (C.c_uint64, [C.c_uint64, C.c_int]),
            "orc_ip_start": (C.c_uint32, [C.c_uint64, C.c_uint32]),
            "orc_ip_move_check": (C.c_int64, [vp, u8p, u32p, C.c_uint32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def large_net_threshold(n, factor=0.01):
    return int(lib().orc_large_net_threshold(n, factor))


def restore_degree_zero(removed, weight, k, part_w, part):
    """Bin packing of the removed vertices; returns (part, part_w)."""
    removed = np.ascontiguousarray(removed, np.uint32)
    weight = np.ascontiguousarray(weight, np.int32)
    pw = np.array(part_w, np.int64)
    pt = np.array(part, np.int32)
    lib().orc_restore_degree_zero(len(removed), removed, len(weight), weight, k, pw, pt)
    return pt, pw


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Hypergraph:
    """Host hypergraph handle of the oracle."""

    def __init__(self, n, noff, pins, vw=None, nw=None, handle=None):
        if handle is not None:
            self.h = handle
        else:
            noff = np.ascontiguousarray(noff, dtype=np.uint64)
            pins = np.ascontiguousarray(pins, dtype=np.uint32)
            self._vw = None if vw is None else np.ascontiguousarray(vw, dtype=np.int32)
            self._nw = None if nw is None else np.ascontiguousarray(nw, dtype=np.int32)
            self.h = lib().orc_hg_new(n, len(noff) - 1, noff, pins, _ptr(self._vw), _ptr(self._nw))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_hg_free(self.h)
            self.h = None

    def sizes(self):
        n, m, p, me, tw = C.c_uint32(), C.c_uint32(), C.c_uint64(), C.c_uint32(), C.c_int64()
        lib().orc_hg_sizes(self.h, C.byref(n), C.byref(m), C.byref(p), C.byref(me), C.byref(tw))
        return n.value, m.value, p.value, me.value, tw.value

    def arrays(self):
        n, m, p, _, _ = self.sizes()
        noff = np.zeros(m + 1, np.uint64); pins = np.zeros(p, np.uint32)
        voff = np.zeros(n + 1, np.uint64); inc = np.zeros(p, np.uint32)
        vw = np.zeros(n, np.int32); nw = np.zeros(m, np.int32); comm = np.zeros(n, np.int32)
        lib().orc_hg_get(self.h, noff, pins, voff, inc, vw, nw, comm)
        return dict(noff=noff, pins=pins, voff=voff, inc=inc, vw=vw, nw=nw, comm=comm)

    def set_communities(self, comm):
        lib().orc_hg_set_comm(self.h, np.ascontiguousarray(comm, dtype=np.int32))

    def phg_init(self, k, part):
        n, m, _, _, _ = self.sizes()
        cw = (k + 63) // 64
        pw = np.zeros(k, np.int64); pc = np.zeros(m * k, np.uint32); conn = np.zeros(max(m * cw, 1), np.uint64)
        km1 = lib().orc_phg_init(self.h, k, np.ascontiguousarray(part, dtype=np.int32), pw, pc, conn)
        return dict(km1=km1, part_w=pw, pin_counts=pc.reshape(m, k), conn=conn[:m * cw].reshape(m, cw))

    def cut(self, k, part):
        return lib().orc_cut(self.h, k, np.ascontiguousarray(part, dtype=np.int32))

    def border(self, k, part):
        n = self.sizes()[0]
        out = np.zeros(n, np.uint8)
        lib().orc_border(self.h, k, np.ascontiguousarray(part, dtype=np.int32), out)
        return out

    def max_gain_move(self, k, part, u, max_w):
        to, gain = C.c_int32(), C.c_int32()
        lib().orc_max_gain_move(self.h, k, np.ascontiguousarray(part, dtype=np.int32), u,
                                np.ascontiguousarray(max_w, dtype=np.int64), C.byref(to), C.byref(gain))
        return to.value, gain.value

    def lp_refine(self, k, part, max_w, perfect_w, seed=0, max_rounds=5, sub_rounds=4, sync=True, threads=1,
                  disable_rand=False, rebalancing=True, nodes=None):
        part = np.array(part, dtype=np.int32)
        st = np.zeros(5, np.int64)
        nd = np.zeros(1, np.uint32) if nodes is None else np.ascontiguousarray(nodes, dtype=np.uint32)
        km1 = lib().orc_lp_refine(self.h, k, part, np.ascontiguousarray(max_w, dtype=np.int64),
                                  np.ascontiguousarray(perfect_w, dtype=np.int64), seed, max_rounds, sub_rounds,
                                  int(sync), threads, int(disable_rand), int(rebalancing), nd,
                                  0 if nodes is None else len(nd), st)
        return part, km1, dict(delta=int(st[0]), moves=int(st[1]), reverts=int(st[2]), visits=int(st[3]),
                               rounds=int(st[4]))

    def fm_refine(self, k, part, max_w, seed=0, rounds=10, seeds=25, pq=64, max_moves=64, ignore=1000,
                  delta_nets=256, expand_cap=1024, sub_rounds=8):
        """Synchronous multitry localized k-way FM; returns (parts, km1, stats)."""
        part = np.array(part, dtype=np.int32)
        st = np.zeros(4, np.int64)
        km1 = lib().orc_fm_refine(self.h, k, part, np.ascontiguousarray(max_w, dtype=np.int64), seed, rounds, seeds,
                                  pq, max_moves, ignore, delta_nets, expand_cap, sub_rounds, st)
        return part, km1, dict(improvement=int(st[0]), moves=int(st[1]), rounds=int(st[2]), searches=int(st[3]))

    def fm_recalculate(self, k, part, v, frm, to):
        """recalculateGains: exact gain of every move of the sequence applied in order."""
        v = np.ascontiguousarray(v, dtype=np.uint32)
        g = np.zeros(max(len(v), 1), np.int64)
        lib().orc_fm_recalculate(self.h, k, np.ascontiguousarray(part, dtype=np.int32), len(v), v,
                                 np.ascontiguousarray(frm, dtype=np.int32), np.ascontiguousarray(to, dtype=np.int32), g)
        return g[:len(v)]

    def fm_best_prefix(self, k, part, v, frm, to, gain, max_w):
        """Best balanced prefix of a move sequence: (length, gain)."""
        v = np.ascontiguousarray(v, dtype=np.uint32)
        bg = C.c_int64()
        n = lib().orc_fm_best_prefix(self.h, k, np.ascontiguousarray(part, dtype=np.int32), len(v), v,
                                     np.ascontiguousarray(frm, dtype=np.int32), np.ascontiguousarray(to, dtype=np.int32),
                                     np.ascontiguousarray(gain, dtype=np.int64),
                                     np.ascontiguousarray(max_w, dtype=np.int64), C.byref(bg))
        return int(n), bg.value

    def rate_sync(self, pass_nr, sub_round, cid, cw, matched, proposers, max_node_weight, ignore=1000,
                  sampling=200000, seed=0, prefer_unmatched=True):
        """The rating step of one synchronous clustering sub-round on the given
        state: (target per proposer, 0xffffffff for none; 36.28 fixed-point
        rating per proposer)."""
        pr = np.ascontiguousarray(proposers, dtype=np.uint32)
        t = np.zeros(max(len(pr), 1), np.uint32)
        r = np.zeros(max(len(pr), 1), np.uint64)
        lib().orc_rate_sync(self.h, pass_nr, sub_round, max_node_weight, ignore, sampling, seed,
                            int(prefer_unmatched), np.ascontiguousarray(cid, dtype=np.uint32),
                            np.ascontiguousarray(cw, dtype=np.int64), np.ascontiguousarray(matched, dtype=np.uint8),
                            pr, len(pr), t, r)
        return t[:len(pr)], r[:len(pr)]

    def rating_sample_ref(self, vertices, ignore=1000, threads=1):
        """Reference-semantics rating of the given vertices on the level-0
        state (OpenMP threads, no matching): pin visits."""
        v = np.ascontiguousarray(vertices, dtype=np.uint32)
        return int(lib().orc_rating_sample_ref(self.h, v, len(v), ignore, threads))

    def cluster_pass(self, sync, pass_nr, contraction_limit, max_node_weight, max_shrink=2.5, ignore=1000,
                     sampling=200000, seed=0, sub_rounds=16, disable_rand=False, prefer_unmatched=True,
                     with_visits=False):
        n = self.sizes()[0]
        out = np.zeros(n, np.uint32)
        visits = np.zeros(1, np.int64)
        after = lib().orc_cluster_pass(self.h, int(sync), pass_nr, contraction_limit, max_node_weight, max_shrink,
                                       ignore, sampling, seed, sub_rounds, int(disable_rand), int(prefer_unmatched),
                                       out, visits)
        return (out, after, int(visits[0])) if with_visits else (out, after)

    def contract(self, cluster):
        n = self.sizes()[0]
        cmap = np.zeros(n, np.uint32)
        h = lib().orc_hg_contract(self.h, np.ascontiguousarray(cluster, dtype=np.uint32), cmap)
        return Hypergraph(0, None, None, handle=h), cmap

    def coarsen(self, sync, contraction_limit, max_node_weight, min_shrink=1.01, max_shrink=2.5, ignore=1000,
                sampling=200000, seed=0, sub_rounds=16, disable_rand=False, prefer_unmatched=True):
        out = C.c_void_p()
        levels = lib().orc_coarsen(self.h, int(sync), contraction_limit, max_node_weight, min_shrink, max_shrink,
                                   ignore, sampling, seed, sub_rounds, int(disable_rand), int(prefer_unmatched),
                                   C.byref(out))
        return levels, (Hypergraph(0, None, None, handle=out.value) if out.value else None)

    def rebalance(self, k, part, max_w):
        """km1 rebalancer (shared host code): (new part, km1 change, moves,
        balanced)."""
        p = np.array(part, np.int32)
        mv, bal = C.c_int64(), C.c_int32()
        d = lib().orc_rebalance(self.h, k, p, np.ascontiguousarray(max_w, np.int64), C.byref(mv), C.byref(bal))
        return p, int(d), mv.value, bool(bal.value)

    def reduce(self, contraction_limit, net_threshold):
        """Degree-zero vertex and large-net removal: (reduced hypergraph,
        vertex map with 0xffffffff for removed vertices, removed vertices,
        removed nets)."""
        n = self.sizes()[0]
        vmap = np.zeros(n, np.uint32)
        rem = np.zeros(2, np.uint32)
        h = lib().orc_reduce(self.h, contraction_limit, net_threshold, vmap, rem)
        return Hypergraph(0, None, None, handle=h), vmap, int(rem[0]), int(rem[1])

    def initial_partition(self, k, max_w, seed, runs=20, threads=4):
        """Recursive-bisection initial partitioning (the device's form):
        (block per vertex, km1)."""
        n = self.sizes()[0]
        part = np.zeros(max(n, 1), np.int32)
        km1 = lib().orc_initial_partition(self.h, k, np.ascontiguousarray(max_w, np.int64), seed, runs, threads,
                                          part)
        return part[:n], int(km1)

    def ip_bisect_run(self, max0, max1, target0, stream, r, passes=4):
        """One run of one bisection: (side per vertex, cut, balanced, side-0 weight)."""
        n = self.sizes()[0]
        side = np.zeros(max(n, 1), np.uint8)
        out = np.zeros(3, np.int64)
        lib().orc_ip_bisect_run(self.h, max0, max1, target0, stream, r, passes, side, out)
        return side[:n], int(out[0]), bool(out[1]), int(out[2])

    def ip_move_check(self, side, moves):
        """Pin counts and gains that differ from a recount after each of the
        moves, applied with the initial partitioner's incremental upkeep."""
        sd = np.ascontiguousarray(side, np.uint8)
        mv = np.ascontiguousarray(moves, np.uint32)
        return int(lib().orc_ip_move_check(self.h, sd, mv, len(mv)))

    def partition(self, k, eps, seed, sync, threads=1, fm_rounds=0):
        """The fast-preset pipeline (LP at every level, then FM when
        fm_rounds > 0), with the device pipeline's defaults."""
        n = self.sizes()[0]
        part = np.zeros(n, np.int32)
        times = np.zeros(5, np.float64)
        cnt = np.zeros(7, np.int64)
        km1 = lib().orc_partition(self.h, k, eps, seed, int(sync), threads, fm_rounds, part, times, cnt)
        return part, km1, dict(louvain_s=times[0], coarsen_s=times[1], ip_s=times[2], refine_s=times[3],
                               total_s=times[4], louvain_arcs=int(cnt[0]), lp_visits=int(cnt[1]),
                               levels=int(cnt[2]), rating_visits=int(cnt[3]), removed_vertices=int(cnt[4]),
                               removed_nets=int(cnt[5]), fm_improvement=int(cnt[6]))


class Graph:
    UNIFORM, NON_UNIFORM, DEGREE = 0, 1, 2

    def __init__(self, hg=None, weight=0, handle=None, ranked=False):
        """ranked: net nodes numbered by locality rank, as the device path does."""
        self.h = handle if handle is not None else lib().orc_graph_new(hg.h, weight, int(ranked))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_graph_free(self.h)
            self.h = None

    def sizes(self):
        n, a, tv, md = C.c_uint32(), C.c_uint64(), C.c_double(), C.c_uint64()
        lib().orc_graph_sizes(self.h, C.byref(n), C.byref(a), C.byref(tv), C.byref(md))
        return n.value, a.value, tv.value, md.value

    def arrays(self):
        n, a, _, _ = self.sizes()
        idx = np.zeros(n + 1, np.uint64); head = np.zeros(a, np.uint32)
        w = np.zeros(a, np.float64); vol = np.zeros(n, np.float64)
        lib().orc_graph_get(self.h, idx, head, w, vol)
        return dict(idx=idx, head=head, w=w, vol=vol)

    def contract(self, C_):
        c = np.array(C_, dtype=np.int32)
        h = lib().orc_graph_contract(self.h, c)
        return Graph(handle=h), c

    def modularity(self, C_):
        return lib().orc_modularity(self.h, np.ascontiguousarray(C_, dtype=np.int32))

    def best_cluster(self, C_, u, sampling=200000):
        g = C.c_double()
        to = lib().orc_louvain_best_cluster(self.h, np.ascontiguousarray(C_, dtype=np.int32), u, sampling,
                                            C.byref(g))
        return to, g.value

    def louvain(self, sync, seed=0, max_passes=5, min_frac=0.01, sampling=200000, sub_rounds=8, threads=1,
                disable_rand=False):
        n = self.sizes()[0]
        out = np.zeros(n, np.int32)
        st = np.zeros(4, np.int64)
        lib().orc_louvain(self.h, int(sync), seed, max_passes, min_frac, sampling, sub_rounds, threads,
                          int(disable_rand), out, st)
        return out, dict(arcs=int(st[0]), rounds=int(st[1]), levels=int(st[2]), moves=int(st[3]))

    def louvain_level_sync_trace(self, seed=0, max_passes=5, min_frac=0.01, sampling=200000, sub_rounds=8):
        """Level-0 synchronous local moving with every sub-round recorded:
        (final clustering, [(members, moves u, moves to) per sub-round])."""
        n = self.sizes()[0]
        cap = n * max_passes + 1
        out = np.zeros(max(n, 1), np.int32)
        mem = np.zeros(cap, np.uint32); moff = np.zeros(max_passes * sub_rounds + 2, np.uint64)
        mu = np.zeros(cap, np.uint32); mt = np.zeros(cap, np.int32); voff = np.zeros_like(moff)
        cnt = lib().orc_louvain_level_sync_trace(self.h, seed, max_passes, min_frac, sampling, sub_rounds, out,
                                                 mem, cap, moff, mu, mt, cap, voff, len(moff))
        assert cnt >= 0
        subs = [(mem[moff[i]:moff[i + 1]], mu[voff[i]:voff[i + 1]], mt[voff[i]:voff[i + 1]]) for i in range(cnt)]
        return out[:n], subs

    def louvain_level_sync(self, seed=0, max_passes=5, min_frac=0.01, sampling=200000, sub_rounds=8):
        n = self.sizes()[0]
        out = np.zeros(n, np.int32)
        st = np.zeros(4, np.int64)
        changed = lib().orc_louvain_level_sync(self.h, seed, max_passes, min_frac, sampling, sub_rounds, out, st)
        return out, bool(changed), dict(arcs=int(st[0]), rounds=int(st[1]), moves=int(st[3]))


def ip_perm(key, n):
    """The keyed Feistel permutation of [0, n) the initial partitioner visits
    vertices in."""
    out = np.zeros(max(n, 1), np.uint32)
    lib().orc_ip_perm(key, n, out)
    return out[:n]


def ip_stream(seed, first, k):
    return int(lib().orc_ip_stream(seed, first, k))


def ip_run_key(stream, r):
    return int(lib().orc_ip_run_key(stream, r))


def ip_start(run_key, n):
    return int(lib().orc_ip_start(run_key, n))
