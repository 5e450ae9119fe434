"""ctypes bindings of libmtkahypar_amd.so (the C ABI in include/mtkahypar_amd.h).

Raises at import-use time if the library has not been built; there is no
fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmtkahypar_amd.so")
_lib = None

vp = C.c_void_p
P = C.POINTER


class HypergraphInfo(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_nets", C.c_uint32), ("num_pins", C.c_uint64),
                ("max_net_size", C.c_uint32), ("total_weight", C.c_int64)]


class LpParams(C.Structure):
    _fields_ = [("max_rounds", C.c_int32), ("sub_rounds", C.c_int32), ("activation_threshold", C.c_uint32),
                ("rebalancing", C.c_int32), ("seed", C.c_uint64), ("max_part_weights", P(C.c_int64)),
                ("perfect_part_weights", P(C.c_int64))]


class LpStats(C.Structure):
    _fields_ = [("delta", C.c_int64), ("moves", C.c_int64), ("reverts", C.c_int64), ("visits", C.c_int64),
                ("rounds", C.c_int64), ("device_ms", C.c_double)]


class FmParams(C.Structure):
    _fields_ = [("max_part_weights", P(C.c_int64)), ("rounds", C.c_int32), ("seeds", C.c_int32),
                ("ignore_net_size", C.c_uint32), ("seed", C.c_uint64), ("sub_rounds", C.c_int32)]


class FmStats(C.Structure):
    _fields_ = [("improvement", C.c_int64), ("moves", C.c_int64), ("rounds", C.c_int64), ("searches", C.c_int64),
                ("device_ms", C.c_double)]


class LouvainParams(C.Structure):
    _fields_ = [("edge_weight", C.c_int32), ("max_passes", C.c_int32), ("min_move_fraction", C.c_double),
                ("sampling", C.c_uint64), ("seed", C.c_uint64), ("sub_rounds", C.c_int32)]


class LouvainStats(C.Structure):
    _fields_ = [("arcs_swept", C.c_int64), ("rounds", C.c_int64), ("levels", C.c_int64), ("moves", C.c_int64),
                ("device_ms", C.c_double)]


class CoarsenParams(C.Structure):
    _fields_ = [("contraction_limit", C.c_uint32), ("max_node_weight", C.c_int32), ("min_shrink_factor", C.c_double),
                ("max_shrink_factor", C.c_double), ("ignore_net_size", C.c_uint32), ("sampling", C.c_uint64),
                ("seed", C.c_uint64), ("sub_rounds", C.c_int32)]


class PartitionParams(C.Structure):
    _fields_ = [("k", C.c_int32), ("epsilon", C.c_double), ("seed", C.c_uint64),
                ("use_community_detection", C.c_int32), ("lp_sub_rounds", C.c_int32),
                ("louvain_sub_rounds", C.c_int32), ("coarsening_sub_rounds", C.c_int32), ("ip_runs", C.c_int32),
                ("louvain_edge_weight", C.c_int32), ("louvain_max_passes", C.c_int32),
                ("louvain_min_move_fraction", C.c_double), ("louvain_sampling", C.c_uint64),
                ("min_shrink_factor", C.c_double), ("max_shrink_factor", C.c_double),
                ("max_allowed_weight_multiplier", C.c_double), ("contraction_limit_multiplier", C.c_uint32),
                ("ignore_net_size", C.c_uint32), ("rating_sampling", C.c_uint64), ("lp_enabled", C.c_int32),
                ("lp_max_rounds", C.c_int32), ("lp_activation_threshold", C.c_uint32), ("lp_rebalancing", C.c_int32),
                ("ip_threads", C.c_int32), ("large_net_remove_factor", C.c_double), ("fm_rounds", C.c_int32),
                ("fm_seeds", C.c_int32)]


class PartitionStats(C.Structure):
    _fields_ = [("km1", C.c_int64), ("imbalance", C.c_double), ("levels", C.c_int32), ("louvain_ms", C.c_double),
                ("coarsening_ms", C.c_double), ("initial_ms", C.c_double), ("refinement_ms", C.c_double),
                ("total_ms", C.c_double), ("louvain_arcs", C.c_int64), ("rating_visits", C.c_int64),
                ("lp_visits", C.c_int64), ("removed_vertices", C.c_int64), ("removed_nets", C.c_int64),
                ("fm_ms", C.c_double), ("fm_improvement", C.c_int64)]


class BatchLpParams(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("max_rounds", C.c_int32), ("sub_rounds", C.c_int32),
                ("activation_threshold", C.c_uint32), ("rebalancing", C.c_int32)]


class ServerParams(C.Structure):
    _fields_ = [("device", C.c_int32), ("slots", C.c_int32), ("hbm_budget_bytes", C.c_uint64),
                ("spin_wait", C.c_int32), ("ip_threads", C.c_int32), ("host_threads", C.c_int32)]


class RequestResult(C.Structure):
    _fields_ = [("km1", C.c_int64), ("imbalance", C.c_double), ("levels", C.c_int32), ("slot", C.c_int32),
                ("queue_ms", C.c_double), ("service_ms", C.c_double), ("latency_ms", C.c_double),
                ("estimate_bytes", C.c_uint64), ("peak_bytes", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("max_in_flight", C.c_int32), ("max_reserved_bytes", C.c_uint64), ("budget_bytes", C.c_uint64),
                ("trims", C.c_uint64)]


class PowerlawParams(C.Structure):
    _fields_ = [("num_vertices", C.c_uint32), ("num_nets", C.c_uint32), ("exponent", C.c_double),
                ("min_size", C.c_uint32), ("max_size", C.c_uint32), ("window", C.c_uint32),
                ("local_fraction", C.c_double), ("vertex_weight_min", C.c_int32),
                ("vertex_weight_max", C.c_int32), ("seed", C.c_uint64)]


class RownetParams(C.Structure):
    _fields_ = [("num_rows", C.c_uint32), ("num_cols", C.c_uint32), ("mean_nnz", C.c_double),
                ("max_nnz", C.c_uint32), ("band", C.c_uint32), ("band_fraction", C.c_double),
                ("seed", C.c_uint64)]


# name -> (restype, argtypes); every function returns int status unless noted
SIGNATURES = {
    "mkp_last_error": (C.c_char_p, []),
    "mkp_context_create": (C.c_int, [C.c_int, P(vp)]),
    "mkp_context_destroy": (C.c_int, [vp]),
    "mkp_context_synchronize": (C.c_int, [vp]),
    "mkp_context_set_profiling": (C.c_int, [vp, C.c_int]),
    "mkp_context_reset_profile": (C.c_int, [vp]),
    "mkp_context_get_profile": (C.c_int, [vp, C.c_int, P(C.c_double), P(C.c_double), P(C.c_int64)]),
    "mkp_profile_family_name": (C.c_char_p, [C.c_int]),
    "mkp_context_profile_kernel_count": (C.c_int, [vp, P(C.c_int32)]),
    "mkp_context_profile_kernel": (C.c_int, [vp, C.c_int32, P(C.c_char_p), P(C.c_double), P(C.c_double),
                                             P(C.c_int64)]),
    "mkp_stream_copy": (C.c_int, [vp, C.c_uint64, C.c_int32, P(C.c_double)]),
    "mkp_partition_initialize": (C.c_int, [vp]),
    "mkp_partition_copy_parts": (C.c_int, [vp, vp]),
    "mkp_hypergraph_create": (C.c_int, [vp, C.c_uint32, C.c_uint32, vp, vp, vp, vp, P(vp)]),
    "mkp_hypergraph_destroy": (C.c_int, [vp]),
    "mkp_hypergraph_get_info": (C.c_int, [vp, P(HypergraphInfo)]),
    "mkp_hypergraph_set_communities": (C.c_int, [vp, vp]),
    "mkp_hypergraph_copy_to_host": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "mkp_partition_create": (C.c_int, [vp, C.c_int32, P(vp)]),
    "mkp_partition_destroy": (C.c_int, [vp]),
    "mkp_partition_set_parts": (C.c_int, [vp, vp]),
    "mkp_partition_get_parts": (C.c_int, [vp, vp]),
    "mkp_partition_get_part_weights": (C.c_int, [vp, vp]),
    "mkp_partition_get_pin_counts": (C.c_int, [vp, vp]),
    "mkp_partition_get_connectivity": (C.c_int, [vp, vp]),
    "mkp_partition_get_border": (C.c_int, [vp, vp]),
    "mkp_partition_get_border_noalloc": (C.c_int, [vp, vp]),
    "mkp_partition_km1": (C.c_int, [vp, P(C.c_int64)]),
    "mkp_lp_refine": (C.c_int, [vp, P(LpParams), P(LpStats)]),
    "mkp_fm_refine": (C.c_int, [vp, P(FmParams), P(FmStats)]),
    "mkp_lp_refine_nodes": (C.c_int, [vp, P(LpParams), vp, C.c_uint32, P(LpStats)]),
    "mkp_read_hgr": (C.c_int, [C.c_char_p, P(vp), P(C.c_uint64)]),
    "mkp_write_hgr": (C.c_int, [vp, C.c_char_p]),
    "mkp_write_partition_file": (C.c_int, [vp, C.c_uint32, C.c_char_p]),
    "mkp_louvain": (C.c_int, [vp, P(LouvainParams), vp, P(LouvainStats)]),
    "mkp_louvain_graph_nodes": (C.c_int, [vp, P(C.c_int64)]),
    "mkp_louvain_graph": (C.c_int, [vp, P(LouvainParams), vp, P(LouvainStats)]),
    "mkp_louvain_evaluate": (C.c_int, [vp, P(LouvainParams), vp, vp, vp, P(C.c_double)]),
    "mkp_cluster_pass": (C.c_int, [vp, P(CoarsenParams), C.c_int32, vp, P(C.c_uint32), P(C.c_int64)]),
    "mkp_hypergraph_contract": (C.c_int, [vp, vp, P(vp), vp]),
    "mkp_coarsen": (C.c_int, [vp, P(CoarsenParams), P(vp)]),
    "mkp_hierarchy_num_levels": (C.c_int, [vp, P(C.c_int32)]),
    "mkp_hierarchy_level": (C.c_int, [vp, C.c_int32, P(vp)]),
    "mkp_hierarchy_destroy": (C.c_int, [vp]),
    "mkp_partition_project": (C.c_int, [vp, vp, vp, C.c_int32]),
    "mkp_partition_hypergraph": (C.c_int, [vp, P(PartitionParams), vp, P(PartitionStats)]),
    "mkp_hypergraph_reduce": (C.c_int, [vp, C.c_uint32, C.c_uint32, P(vp), vp, P(C.c_uint32), P(C.c_uint32)]),
    "mkp_rebalance": (C.c_int, [vp, vp, P(C.c_int64), P(C.c_int64)]),
    "mkp_large_net_threshold": (C.c_uint32, [C.c_uint32, C.c_double]),
    "mkp_partition_params_default": (C.c_int, [P(PartitionParams), C.c_int32, C.c_double, C.c_uint64]),
    "mkp_initial_partition": (C.c_int, [vp, C.c_double, C.c_uint64, C.c_int32, P(C.c_int64)]),
    "mkp_ip_bisect_runs": (C.c_int, [vp, C.c_int64, C.c_int64, C.c_int64, C.c_uint64, C.c_int32, C.c_int32, vp, vp]),
    "mkp_generate_powerlaw": (C.c_int, [P(PowerlawParams), P(vp)]),
    "mkp_generate_rownet": (C.c_int, [P(RownetParams), P(vp)]),
    "mkp_host_hypergraph_info": (C.c_int, [vp, P(HypergraphInfo)]),
    "mkp_host_hypergraph_copy": (C.c_int, [vp, vp, vp, vp, vp]),
    "mkp_host_hypergraph_destroy": (C.c_int, [vp]),
    "mkp_hypergraph_create_from_host": (C.c_int, [vp, vp, P(vp)]),
    "mkp_batch_create": (C.c_int, [vp, C.c_uint32, P(vp), P(vp)]),
    "mkp_batch_destroy": (C.c_int, [vp]),
    "mkp_batch_get_info": (C.c_int, [vp, P(C.c_uint32), P(C.c_uint32), P(C.c_uint32), P(C.c_uint64)]),
    "mkp_batch_hypergraph": (C.c_int, [vp, P(vp)]),
    "mkp_batch_louvain": (C.c_int, [vp, P(LouvainParams), vp, vp, vp]),
    "mkp_context_set_profile_filter": (C.c_int, [vp, C.c_char_p]),
    "mkp_context_set_tuning": (C.c_int, [vp, C.c_char_p, C.c_int64]),
    "mkp_batch_partition_copy_parts": (C.c_int, [vp, vp]),
    "mkp_request_estimate": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32, C.c_int32,
                                       P(C.c_uint64)]),
    "mkp_server_create": (C.c_int, [P(ServerParams), P(vp)]),
    "mkp_server_destroy": (C.c_int, [vp]),
    "mkp_server_submit": (C.c_int, [vp, vp, P(PartitionParams), P(C.c_uint64)]),
    "mkp_server_wait": (C.c_int, [vp, C.c_uint64, vp, P(RequestResult)]),
    "mkp_server_get_stats": (C.c_int, [vp, P(ServerStats)]),
    "mkp_host_hypergraph_create": (C.c_int, [C.c_uint32, C.c_uint32, vp, vp, vp, vp, P(vp)]),
    "mkp_batch_partition_create": (C.c_int, [vp, vp, P(vp)]),
    "mkp_batch_partition_destroy": (C.c_int, [vp]),
    "mkp_batch_partition_set_parts": (C.c_int, [vp, vp]),
    "mkp_batch_partition_get_parts": (C.c_int, [vp, vp]),
    "mkp_batch_partition_km1": (C.c_int, [vp, vp]),
    "mkp_batch_lp_refine": (C.c_int, [vp, P(BatchLpParams), vp, vp]),
}


class MkpError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MkpError("libmtkahypar_amd.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != 0:
        raise MkpError("status %d: %s" % (status, lib().mkp_last_error().decode()))


def _p(a):
    return None if a is None else a.ctypes.data_as(vp)


class Context:
    def __init__(self, device=0):
        self.h = vp()
        check(lib().mkp_context_create(device, C.byref(self.h)))

    def close(self):
        if self.h:
            check(lib().mkp_context_destroy(self.h))
            self.h = vp()

    def synchronize(self):
        check(lib().mkp_context_synchronize(self.h))

    FAMILIES = 9

    def set_profiling(self, on):
        check(lib().mkp_context_set_profiling(self.h, int(on)))

    def reset_profile(self):
        check(lib().mkp_context_reset_profile(self.h))

    def set_tuning(self, key, value):
        check(lib().mkp_context_set_tuning(self.h, key.encode(), int(value)))

    def set_profile_filter(self, kernel_name=None):
        """Time only the launches of one kernel (None: every kernel)."""
        check(lib().mkp_context_set_profile_filter(self.h, None if kernel_name is None else kernel_name.encode()))

    def profile(self):
        out = {}
        for f in range(self.FAMILIES):
            ms, by, n = C.c_double(), C.c_double(), C.c_int64()
            check(lib().mkp_context_get_profile(self.h, f, C.byref(ms), C.byref(by), C.byref(n)))
            out[lib().mkp_profile_family_name(f).decode()] = dict(ms=ms.value, bytes=by.value, launches=n.value)
        return out

    def kernel_profile(self):
        """Per-kernel records of single-launch timings: name -> ms, bytes, launches."""
        cnt = C.c_int32()
        check(lib().mkp_context_profile_kernel_count(self.h, C.byref(cnt)))
        out = {}
        for i in range(cnt.value):
            name, ms, by, n = C.c_char_p(), C.c_double(), C.c_double(), C.c_int64()
            check(lib().mkp_context_profile_kernel(self.h, i, C.byref(name), C.byref(ms), C.byref(by), C.byref(n)))
            out[name.value.decode()] = dict(ms=ms.value, bytes=by.value, launches=n.value)
        return out

    def stream_copy(self, nbytes, reps):
        """Mean milliseconds of one device copy of nbytes."""
        ms = C.c_double()
        check(lib().mkp_stream_copy(self.h, nbytes, reps, C.byref(ms)))
        return ms.value


class HostHypergraph:
    """Synthetic input generated on the host."""

    def __init__(self, handle):
        self.h = handle

    @staticmethod
    def powerlaw(num_vertices, num_nets, exponent, min_size, max_size, window=0, local_fraction=0.0,
                 vertex_weight_min=1, vertex_weight_max=1, seed=42):
        prm = PowerlawParams(num_vertices, num_nets, exponent, min_size, max_size, window, local_fraction,
                             vertex_weight_min, vertex_weight_max, seed)
        h = vp()
        check(lib().mkp_generate_powerlaw(C.byref(prm), C.byref(h)))
        return HostHypergraph(h)

    @staticmethod
    def rownet(num_rows, num_cols, mean_nnz, max_nnz, band, band_fraction, seed=42):
        prm = RownetParams(num_rows, num_cols, mean_nnz, max_nnz, band, band_fraction, seed)
        h = vp()
        check(lib().mkp_generate_rownet(C.byref(prm), C.byref(h)))
        return HostHypergraph(h)

    @staticmethod
    def from_arrays(n, noff, pins, vw=None, nw=None):
        noff = np.ascontiguousarray(noff, dtype=np.uint64)
        pins = np.ascontiguousarray(pins, dtype=np.uint32)
        vw = None if vw is None else np.ascontiguousarray(vw, dtype=np.int32)
        nw = None if nw is None else np.ascontiguousarray(nw, dtype=np.int32)
        h = vp()
        check(lib().mkp_host_hypergraph_create(n, len(noff) - 1, _p(noff), _p(pins), _p(vw), _p(nw), C.byref(h)))
        return HostHypergraph(h)

    @staticmethod
    def read_hgr(path):
        """hMetis file; returns (hypergraph, number of dropped single-pin nets)."""
        h, removed = vp(), C.c_uint64()
        check(lib().mkp_read_hgr(str(path).encode(), C.byref(h), C.byref(removed)))
        return HostHypergraph(h), removed.value

    def write_hgr(self, path):
        check(lib().mkp_write_hgr(self.h, str(path).encode()))

    def info(self):
        i = HypergraphInfo()
        check(lib().mkp_host_hypergraph_info(self.h, C.byref(i)))
        return i

    def arrays(self):
        i = self.info()
        noff = np.zeros(i.num_nets + 1, np.uint64)
        pins = np.zeros(max(i.num_pins, 1), np.uint32)
        vw = np.zeros(max(i.num_vertices, 1), np.int32)
        nw = np.zeros(max(i.num_nets, 1), np.int32)
        check(lib().mkp_host_hypergraph_copy(self.h, _p(noff), _p(pins), _p(vw), _p(nw)))
        return dict(n=i.num_vertices, noff=noff, pins=pins[:i.num_pins], vw=vw[:i.num_vertices],
                    nw=nw[:i.num_nets])

    def close(self):
        if self.h:
            check(lib().mkp_host_hypergraph_destroy(self.h))
            self.h = vp()


class Hypergraph:
    def __init__(self, ctx, n=None, noff=None, pins=None, vw=None, nw=None, host=None, handle=None):
        self.ctx = ctx
        self.h = vp()
        if handle is not None:
            self.h = handle
        elif host is not None:
            check(lib().mkp_hypergraph_create_from_host(ctx.h, host.h, C.byref(self.h)))
        else:
            self._noff = np.ascontiguousarray(noff, dtype=np.uint64)
            self._pins = np.ascontiguousarray(pins, dtype=np.uint32)
            self._vw = None if vw is None else np.ascontiguousarray(vw, dtype=np.int32)
            self._nw = None if nw is None else np.ascontiguousarray(nw, dtype=np.int32)
            check(lib().mkp_hypergraph_create(ctx.h, n, len(self._noff) - 1, _p(self._noff), _p(self._pins),
                                              _p(self._vw), _p(self._nw), C.byref(self.h)))

    def info(self):
        i = HypergraphInfo()
        check(lib().mkp_hypergraph_get_info(self.h, C.byref(i)))
        return i

    def set_communities(self, comm):
        c = np.ascontiguousarray(comm, dtype=np.int32)
        check(lib().mkp_hypergraph_set_communities(self.h, _p(c)))

    def ip_bisect_runs(self, max0, max1, target0, stream, runs=20, passes=4):
        """The runs of one initial-partitioning bisection: (sides runs x n,
        results runs x [cut, balanced, side-0 weight, side-1 weight])."""
        n = self.info().num_vertices
        sides = np.zeros((runs, max(n, 1)), np.uint8)
        res = np.zeros((runs, 4), np.int64)
        check(lib().mkp_ip_bisect_runs(self.h, max0, max1, target0, stream, runs, passes, _p(sides), _p(res)))
        return sides[:, :n], res

    UNIFORM, NON_UNIFORM, DEGREE, HYBRID = 0, 1, 2, -1

    def louvain(self, edge_weight=-1, max_passes=5, min_move_fraction=0.01, sampling=200000, seed=0,
                sub_rounds=8, want_communities=True):
        prm = LouvainParams(edge_weight, max_passes, min_move_fraction, sampling, seed, sub_rounds)
        st = LouvainStats()
        out = np.zeros(max(self.info().num_vertices, 1), np.int32) if want_communities else None
        check(lib().mkp_louvain(self.h, C.byref(prm), _p(out), C.byref(st)))
        stats = dict(arcs_swept=st.arcs_swept, rounds=st.rounds, levels=st.levels, moves=st.moves,
                     device_ms=st.device_ms)
        return (out[:self.info().num_vertices] if out is not None else None), stats

    def louvain_graph(self, edge_weight=-1, max_passes=5, min_move_fraction=0.01, sampling=200000, seed=0,
                      sub_rounds=8):
        """Clustering of every node of the Louvain graph (N + E for the star
        expansion, net e at node N + e), as run_parallel_louvain returns it."""
        nodes = C.c_int64()
        check(lib().mkp_louvain_graph_nodes(self.h, C.byref(nodes)))
        prm = LouvainParams(edge_weight, max_passes, min_move_fraction, sampling, seed, sub_rounds)
        st = LouvainStats()
        out = np.zeros(max(nodes.value, 1), np.int32)
        check(lib().mkp_louvain_graph(self.h, C.byref(prm), _p(out), C.byref(st)))
        stats = dict(arcs_swept=st.arcs_swept, rounds=st.rounds, levels=st.levels, moves=st.moves,
                     device_ms=st.device_ms)
        return out[:nodes.value], stats

    def louvain_evaluate(self, node_communities, edge_weight=-1, sampling=200000):
        """Best-cluster decision, FP64 best gain per node of the level-0 graph
        on the given clustering, and its modularity (nothing moves)."""
        prm = LouvainParams(edge_weight, 5, 0.01, sampling, 0, 1)
        cm = np.ascontiguousarray(node_communities, dtype=np.int32)
        tg = np.zeros(max(len(cm), 1), np.int32)
        gn = np.zeros(max(len(cm), 1), np.float64)
        q = C.c_double()
        check(lib().mkp_louvain_evaluate(self.h, C.byref(prm), _p(cm), _p(tg), _p(gn), C.byref(q)))
        return tg[:len(cm)], gn[:len(cm)], q.value

    @staticmethod
    def coarsen_params(contraction_limit, max_node_weight, min_shrink=1.01, max_shrink=2.5, ignore=1000,
                       sampling=200000, seed=0, sub_rounds=16):
        return CoarsenParams(contraction_limit, max_node_weight, min_shrink, max_shrink, ignore, sampling, seed,
                             sub_rounds)

    def cluster_pass(self, pass_nr, **kw):
        prm = self.coarsen_params(**kw)
        n = self.info().num_vertices
        out = np.zeros(max(n, 1), np.uint32)
        after, visits = C.c_uint32(), C.c_int64()
        check(lib().mkp_cluster_pass(self.h, C.byref(prm), pass_nr, _p(out), C.byref(after), C.byref(visits)))
        return out[:n], after.value, visits.value

    def contract(self, clusters):
        cl = np.ascontiguousarray(clusters, dtype=np.uint32)
        n = self.info().num_vertices
        cmap = np.zeros(max(n, 1), np.uint32)
        h = vp()
        check(lib().mkp_hypergraph_contract(self.h, _p(cl), C.byref(h), _p(cmap)))
        return Hypergraph(self.ctx, handle=h), cmap[:n]

    def coarsen(self, **kw):
        prm = self.coarsen_params(**kw)
        h = vp()
        check(lib().mkp_coarsen(self.h, C.byref(prm), C.byref(h)))
        return Hierarchy(self.ctx, h, self)

    def partition(self, k, epsilon, seed, **overrides):
        """Multilevel partition with the fast preset; keyword arguments
        override fields of the parameter struct (e.g. lp_sub_rounds=8)."""
        prm = PartitionParams()
        check(lib().mkp_partition_params_default(C.byref(prm), k, epsilon, seed))
        for key, val in overrides.items():
            if key not in dict(PartitionParams._fields_):
                raise KeyError(key)
            setattr(prm, key, int(val) if isinstance(val, bool) else val)
        st = PartitionStats()
        out = np.zeros(max(self.info().num_vertices, 1), np.int32)
        check(lib().mkp_partition_hypergraph(self.h, C.byref(prm), _p(out), C.byref(st)))
        stats = {f: getattr(st, f) for f, _ in PartitionStats._fields_}
        return out[:self.info().num_vertices], stats

    def reduce(self, contraction_limit, net_threshold):
        """Degree-zero vertex and large-net removal: (reduced Hypergraph or
        None when nothing was removed, vertex map with 0xffffffff for removed
        vertices, removed vertices, removed nets)."""
        n = self.info().num_vertices
        out = vp()
        vmap = np.zeros(max(n, 1), np.uint32)
        rv, rn = C.c_uint32(), C.c_uint32()
        check(lib().mkp_hypergraph_reduce(self.h, contraction_limit, net_threshold, C.byref(out), _p(vmap),
                                          C.byref(rv), C.byref(rn)))
        red = Hypergraph(self.ctx, handle=out) if out.value else None
        return red, vmap[:n], rv.value, rn.value

    def arrays(self):
        i = self.info()
        noff = np.zeros(i.num_nets + 1, np.uint64); pins = np.zeros(max(i.num_pins, 1), np.uint32)
        voff = np.zeros(i.num_vertices + 1, np.uint64); inc = np.zeros(max(i.num_pins, 1), np.uint32)
        vw = np.zeros(max(i.num_vertices, 1), np.int32); nw = np.zeros(max(i.num_nets, 1), np.int32)
        comm = np.zeros(max(i.num_vertices, 1), np.int32)
        check(lib().mkp_hypergraph_copy_to_host(self.h, _p(noff), _p(pins), _p(voff), _p(inc), _p(vw), _p(nw),
                                                _p(comm)))
        p, n, m = i.num_pins, i.num_vertices, i.num_nets
        return dict(noff=noff, pins=pins[:p], voff=voff, inc=inc[:p], vw=vw[:n], nw=nw[:m], comm=comm[:n])

    def close(self):
        if self.h:
            check(lib().mkp_hypergraph_destroy(self.h))
            self.h = vp()


class Hierarchy:
    """Level hierarchy of a coarsening run (level 0 is the input)."""

    def __init__(self, ctx, handle, top):
        self.ctx, self.h, self.top = ctx, handle, top

    def num_levels(self):
        v = C.c_int32()
        check(lib().mkp_hierarchy_num_levels(self.h, C.byref(v)))
        return v.value

    def level(self, i):
        h = vp()
        check(lib().mkp_hierarchy_level(self.h, i, C.byref(h)))
        hg = Hypergraph(self.ctx, handle=h)
        hg.borrowed = True
        return hg

    def project(self, fine, coarse, level):
        check(lib().mkp_partition_project(fine.h, coarse.h, self.h, level))

    def close(self):
        if self.h:
            check(lib().mkp_hierarchy_destroy(self.h))
            self.h = vp()


class Partition:
    def __init__(self, hg, k):
        self.hg, self.k = hg, k
        self.h = vp()
        check(lib().mkp_partition_create(hg.h, k, C.byref(self.h)))

    def set_parts(self, parts):
        a = np.ascontiguousarray(parts, dtype=np.int32)
        check(lib().mkp_partition_set_parts(self.h, _p(a)))

    def initialize(self):
        check(lib().mkp_partition_initialize(self.h))

    def copy_parts_from(self, src):
        check(lib().mkp_partition_copy_parts(self.h, src.h))

    def parts(self):
        out = np.zeros(self.hg.info().num_vertices, np.int32)
        check(lib().mkp_partition_get_parts(self.h, _p(out)))
        return out

    def part_weights(self):
        out = np.zeros(self.k, np.int64)
        check(lib().mkp_partition_get_part_weights(self.h, _p(out)))
        return out

    def pin_counts(self):
        m = self.hg.info().num_nets
        out = np.zeros(max(m * self.k, 1), np.uint32)
        check(lib().mkp_partition_get_pin_counts(self.h, _p(out)))
        return out[:m * self.k].reshape(m, self.k)

    def connectivity(self):
        m = self.hg.info().num_nets
        cw = (self.k + 63) // 64
        out = np.zeros(max(m * cw, 1), np.uint64)
        check(lib().mkp_partition_get_connectivity(self.h, _p(out)))
        return out[:m * cw].reshape(m, cw)

    def border(self):
        out = np.zeros(max(self.hg.info().num_vertices, 1), np.uint8)
        check(lib().mkp_partition_get_border(self.h, _p(out)))
        return out[:self.hg.info().num_vertices]

    def border_noalloc(self):
        """border() through the entry point that allocates no device memory."""
        out = np.zeros(max(self.hg.info().num_vertices, 1), np.uint8)
        check(lib().mkp_partition_get_border_noalloc(self.h, _p(out)))
        return out[:self.hg.info().num_vertices]

    def km1(self):
        v = C.c_int64()
        check(lib().mkp_partition_km1(self.h, C.byref(v)))
        return v.value

    def initial_partition(self, epsilon, seed, runs=20):
        """Recursive-bisection initial partitioning on the device; returns km1."""
        v = C.c_int64()
        check(lib().mkp_initial_partition(self.h, epsilon, seed, runs, C.byref(v)))
        return v.value

    def rebalance(self, max_w):
        """km1 rebalancer: (km1 change, moves)."""
        mw = np.ascontiguousarray(max_w, dtype=np.int64)
        d, mv = C.c_int64(), C.c_int64()
        check(lib().mkp_rebalance(self.h, _p(mw), C.byref(d), C.byref(mv)))
        return d.value, mv.value

    def lp_refine(self, max_w, perfect_w, seed=0, max_rounds=5, sub_rounds=4, activation_threshold=100,
                  rebalancing=True, nodes=None):
        mw = np.ascontiguousarray(max_w, dtype=np.int64)
        pw = np.ascontiguousarray(perfect_w, dtype=np.int64)
        prm = LpParams(max_rounds, sub_rounds, activation_threshold, int(rebalancing), seed,
                       mw.ctypes.data_as(P(C.c_int64)), pw.ctypes.data_as(P(C.c_int64)))
        st = LpStats()
        if nodes is None:
            check(lib().mkp_lp_refine(self.h, C.byref(prm), C.byref(st)))
        else:
            nd = np.ascontiguousarray(nodes, dtype=np.uint32)
            check(lib().mkp_lp_refine_nodes(self.h, C.byref(prm), _p(nd), len(nd), C.byref(st)))
        return dict(delta=st.delta, moves=st.moves, reverts=st.reverts, visits=st.visits, rounds=st.rounds,
                    device_ms=st.device_ms)

    def fm_refine(self, max_w, seed=0, rounds=10, seeds=25, ignore=1000, sub_rounds=8):
        """Multitry localized k-way FM (synchronous form)."""
        mw = np.ascontiguousarray(max_w, dtype=np.int64)
        prm = FmParams(mw.ctypes.data_as(P(C.c_int64)), rounds, seeds, ignore, seed, sub_rounds)
        st = FmStats()
        check(lib().mkp_fm_refine(self.h, C.byref(prm), C.byref(st)))
        return dict(improvement=st.improvement, moves=st.moves, rounds=st.rounds, searches=st.searches,
                    device_ms=st.device_ms)

    def close(self):
        if self.h:
            check(lib().mkp_partition_destroy(self.h))
            self.h = vp()


class Batch:
    """Many hypergraphs resident as one disjoint union (the batched form)."""

    def __init__(self, ctx, hosts):
        self.ctx = ctx
        arr = (vp * len(hosts))(*[h.h for h in hosts])
        self.h = vp()
        check(lib().mkp_batch_create(ctx.h, len(hosts), arr, C.byref(self.h)))

    def info(self):
        m, n, e, p = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(lib().mkp_batch_get_info(self.h, C.byref(m), C.byref(n), C.byref(e), C.byref(p)))
        return dict(members=m.value, vertices=n.value, nets=e.value, pins=p.value)

    def hypergraph(self):
        h = vp()
        check(lib().mkp_batch_hypergraph(self.h, C.byref(h)))
        hg = Hypergraph(self.ctx, handle=h)
        hg.borrowed = True
        return hg

    def louvain(self, seeds=None, edge_weight=-1, max_passes=5, min_move_fraction=0.01, sampling=200000, seed=0,
                sub_rounds=8, want_communities=True):
        """Louvain on every member; returns (member-local communities of all
        vertices or None, per-member stats)."""
        prm = LouvainParams(edge_weight, max_passes, min_move_fraction, sampling, seed, sub_rounds)
        info = self.info()
        G = info["members"]
        st = (LouvainStats * G)()
        sd = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.uint64)
        out = np.zeros(max(info["vertices"], 1), np.int32) if want_communities else None
        check(lib().mkp_batch_louvain(self.h, C.byref(prm), _p(sd), _p(out), C.cast(st, vp)))
        stats = [dict(arcs_swept=x.arcs_swept, rounds=x.rounds, levels=x.levels, moves=x.moves,
                      device_ms=x.device_ms) for x in st]
        return (out[:info["vertices"]] if out is not None else None), stats

    def close(self):
        if self.h:
            check(lib().mkp_batch_destroy(self.h))
            self.h = vp()


class BatchPartition:
    def __init__(self, batch, k):
        self.batch = batch
        self.k = np.ascontiguousarray(k, dtype=np.int32)
        self.h = vp()
        check(lib().mkp_batch_partition_create(batch.h, _p(self.k), C.byref(self.h)))

    def set_parts(self, parts):
        a = np.ascontiguousarray(parts, dtype=np.int32)
        check(lib().mkp_batch_partition_set_parts(self.h, _p(a)))

    def copy_parts_from(self, src):
        check(lib().mkp_batch_partition_copy_parts(self.h, src.h))

    def parts(self):
        out = np.zeros(max(self.batch.info()["vertices"], 1), np.int32)
        check(lib().mkp_batch_partition_get_parts(self.h, _p(out)))
        return out[:self.batch.info()["vertices"]]

    def km1(self):
        out = np.zeros(len(self.k), np.int64)
        check(lib().mkp_batch_partition_km1(self.h, _p(out)))
        return out

    def lp_refine(self, seeds, epsilon=0.03, max_rounds=5, sub_rounds=4, activation_threshold=100, rebalancing=True):
        prm = BatchLpParams(epsilon, max_rounds, sub_rounds, activation_threshold, int(rebalancing))
        sd = np.ascontiguousarray(seeds, dtype=np.uint64)
        st = (LpStats * len(self.k))()
        check(lib().mkp_batch_lp_refine(self.h, C.byref(prm), _p(sd), C.cast(st, vp)))
        return [dict(delta=x.delta, moves=x.moves, reverts=x.reverts, visits=x.visits, rounds=x.rounds) for x in st]

    def close(self):
        if self.h:
            check(lib().mkp_batch_partition_destroy(self.h))
            self.h = vp()


def large_net_threshold(num_vertices, factor=0.01):
    return int(lib().mkp_large_net_threshold(num_vertices, factor))


def partition_params(k, epsilon, seed, **overrides):
    """Fast-preset parameter struct with keyword overrides."""
    prm = PartitionParams()
    check(lib().mkp_partition_params_default(C.byref(prm), k, epsilon, seed))
    for key, val in overrides.items():
        if key not in dict(PartitionParams._fields_):
            raise KeyError(key)
        setattr(prm, key, int(val) if isinstance(val, bool) else val)
    return prm


def request_estimate(num_vertices, num_nets, num_pins, max_net_size, k):
    out = C.c_uint64()
    check(lib().mkp_request_estimate(num_vertices, num_nets, num_pins, max_net_size, k, C.byref(out)))
    return out.value


class Server:
    """Request loop of the served form: submit() queues a partition request,
    wait() returns its block ids and timings."""

    def __init__(self, device=0, slots=4, hbm_budget_bytes=0, spin_wait=False, ip_threads=0, host_threads=0):
        prm = ServerParams(device, slots, hbm_budget_bytes, 1 if spin_wait else 0, ip_threads, host_threads)
        self.h = vp()
        check(lib().mkp_server_create(C.byref(prm), C.byref(self.h)))
        self._hosts = {}

    def submit(self, host, k, epsilon, seed, **overrides):
        prm = partition_params(k, epsilon, seed, **overrides)
        t = C.c_uint64()
        check(lib().mkp_server_submit(self.h, host.h, C.byref(prm), C.byref(t)))
        self._hosts[t.value] = host  # alive until waited for
        return t.value

    def wait(self, ticket, want_parts=True):
        host = self._hosts.pop(ticket)
        n = host.info().num_vertices
        out = np.zeros(max(n, 1), np.int32) if want_parts else None
        r = RequestResult()
        check(lib().mkp_server_wait(self.h, ticket, _p(out), C.byref(r)))
        res = {f: getattr(r, f) for f, _ in RequestResult._fields_}
        return (out[:n] if out is not None else None), res

    def stats(self):
        st = ServerStats()
        check(lib().mkp_server_get_stats(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            check(lib().mkp_server_destroy(self.h))
            self.h = vp()
