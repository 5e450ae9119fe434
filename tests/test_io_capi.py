"""hMetis reader/writer and the host-only parts of the C library interface
(libkahypar.h: context, ini presets, read_hypergraph_from_file).  CPU only."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import mtk_amd
from helpers import GOLDEN, read_hgr, vectors

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def host_arrays(path):
    h, removed = mtk_amd.HostHypergraph.read_hgr(path)
    a = h.arrays()
    h.close()
    return a, removed


@pytest.mark.parametrize("name", ["karate_club.graph.hgr", "contracted_unweighted_ibm01.hgr"])
def test_reader_matches_python_reader_on_fixtures(name):
    path = os.path.join(GOLDEN, name)
    n, noff, pins, vw, nw = read_hgr(path)
    a, _ = host_arrays(path)
    assert len(a["vw"]) == n
    assert np.array_equal(a["noff"], noff) and np.array_equal(a["pins"], pins)
    assert np.array_equal(a["vw"], vw) and np.array_equal(a["nw"], nw)


def test_reader_formats_comments_and_single_pin_nets(tmp_path):
    # type 11: net weights first on each net line, vertex weights at the end
    text = "% comment\n5 4 11\n% inside\n3 1 2\n7 4\n2 1 3 4\n\n5 2 3\n9\n8\n7\n6\n"
    p = tmp_path / "a.hgr"
    p.write_text(text)
    a, removed = host_arrays(str(p))
    assert removed == 2  # "7 4" (one pin) and the empty line
    assert a["noff"].tolist() == [0, 2, 5, 7]
    assert a["pins"].tolist() == [0, 1, 0, 2, 3, 1, 2]
    assert a["nw"].tolist() == [3, 2, 5]
    assert a["vw"].tolist() == [9, 8, 7, 6]
    # type 10: vertex weights only
    p.write_text("2 3 10\n1 2\n2 3\n4\n5\n6\n")
    a, removed = host_arrays(str(p))
    assert removed == 0 and a["nw"].tolist() == [1, 1] and a["vw"].tolist() == [4, 5, 6]
    # type 1 and unweighted
    p.write_text("2 3 1\n4 1 2\n5 2 3\n")
    assert host_arrays(str(p))[0]["nw"].tolist() == [4, 5]
    p.write_text("2 3\n1 2 3\n3\n")
    a, removed = host_arrays(str(p))
    assert removed == 1 and a["pins"].tolist() == [0, 1, 2]


@pytest.mark.parametrize("text,msg", [("2 3\n1 4\n1 2\n", "outside"), ("3 3\n1 2\n", "net lines"),
                                      ("1 3 7\n1 2\n", "type"), ("1 2 10\n1 2\n5\n", "vertex weights"),
                                      ("1 3\n1 x\n", "integer"), ("1 3\n1 2x\n", "unexpected"),
                                      ("", "header")])
def test_reader_rejects_malformed_files(tmp_path, text, msg):
    p = tmp_path / "bad.hgr"
    p.write_text(text)
    with pytest.raises(mtk_amd.MkpError, match=msg):
        mtk_amd.HostHypergraph.read_hgr(str(p))


def test_writer_round_trip(tmp_path):
    h = mtk_amd.HostHypergraph.powerlaw(500, 700, 2.5, 2, 30, window=64, local_fraction=0.8,
                                        vertex_weight_min=1, vertex_weight_max=5, seed=3)
    a = h.arrays()
    p = tmp_path / "w.hgr"
    h.write_hgr(str(p))
    h.close()
    b, removed = host_arrays(str(p))
    assert removed == 0
    for key in ("noff", "pins", "vw", "nw"):
        assert np.array_equal(a[key], b[key]), key


# ------------------------------------------------------- C library interface
def capi():
    L = C.CDLL(mtk_amd.LIB_PATH)
    L.mt_kahypar_context_new.restype = C.c_void_p
    L.mt_kahypar_context_free.argtypes = [C.c_void_p]
    L.mt_kahypar_configure_context_from_file.argtypes = [C.c_void_p, C.c_char_p]
    L.mt_kahypar_amd_set_error_mode.argtypes = [C.c_int]
    L.mt_kahypar_read_hypergraph_from_file.argtypes = [C.c_char_p, C.POINTER(C.c_ulong), C.POINTER(C.c_ulong),
                                                       C.POINTER(C.POINTER(C.c_size_t)),
                                                       C.POINTER(C.POINTER(C.c_ulong)),
                                                       C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.POINTER(C.c_int))]
    L.mt_kahypar_amd_get_params.argtypes = [C.c_void_p, C.POINTER(mtk_amd.PartitionParams)]
    L.mt_kahypar_amd_get_params.restype = C.c_int
    L.mkp_last_error.restype = C.c_char_p
    return L


def test_capi_read_hypergraph_from_file():
    L = capi()
    path = os.path.join(GOLDEN, "contracted_unweighted_ibm01.hgr")
    nv, ne = C.c_ulong(), C.c_ulong()
    idx, pins = C.POINTER(C.c_size_t)(), C.POINTER(C.c_ulong)()
    ew, vw = C.POINTER(C.c_int)(), C.POINTER(C.c_int)()
    L.mt_kahypar_read_hypergraph_from_file(path.encode(), C.byref(nv), C.byref(ne), C.byref(idx), C.byref(pins),
                                           C.byref(ew), C.byref(vw))
    n, noff, rpins, rvw, rnw = read_hgr(path)
    assert (nv.value, ne.value) == (n, len(rnw))
    assert [idx[i] for i in range(ne.value + 1)] == noff.tolist()
    assert [pins[i] for i in range(len(rpins))] == rpins.tolist()
    assert [ew[i] for i in range(ne.value)] == rnw.tolist()
    assert [vw[i] for i in range(nv.value)] == rvw.tolist()


PRESET = """# fast preset values read by this path
maxnet-ignore=1000
p-enable-community-detection=true
p-louvain-edge-weight-function=hybrid
p-max-louvain-pass-iterations=5
p-louvain-min-vertex-move-fraction=0.01
p-vertex-degree-sampling-threshold=200000
c-type=multilevel_coarsener
c-min-shrink-factor=1.01
c-max-shrink-factor=2.5
c-s=1
c-t=160
c-vertex-degree-sampling-threshold=200000
i-runs=20
r-lp-type=label_propagation_km1
r-lp-maximum-iterations=5
r-lp-rebalancing=true
r-lp-he-size-activation-threshold=100
# keys of components outside this path are accepted
r-fm-type=fm_gain_cache
s-num-threads=8
"""


def test_capi_context_ini(tmp_path):
    L = capi()
    L.mt_kahypar_amd_set_error_mode(1)
    ctx = L.mt_kahypar_context_new()
    default = mtk_amd.PartitionParams()
    assert L.mt_kahypar_amd_get_params(ctx, C.byref(default)) == 0
    p = tmp_path / "fast.ini"
    p.write_text(PRESET)
    L.mt_kahypar_configure_context_from_file(ctx, str(p).encode())
    got = mtk_amd.PartitionParams()
    L.mt_kahypar_amd_get_params(ctx, C.byref(got))
    # the preset equals the defaults, except that it switches FM on (the
    # defaults leave it off): r-fm-type=fm_gain_cache, 10 rounds, 25 seeds
    for f, _ in mtk_amd.PartitionParams._fields_:
        if f != "fm_rounds":
            assert getattr(got, f) == getattr(default, f), f
    assert (default.fm_rounds, got.fm_rounds, got.fm_seeds) == (0, 10, 25)
    p.write_text("c-t=80\np-louvain-edge-weight-function=degree\nr-lp-type=do_nothing\ni-runs=3\n"
                 "p-enable-community-detection=false\nc-max-shrink-factor=3.0\n")
    L.mt_kahypar_configure_context_from_file(ctx, str(p).encode())
    L.mt_kahypar_amd_get_params(ctx, C.byref(got))
    assert (got.contraction_limit_multiplier, got.louvain_edge_weight, got.lp_enabled, got.ip_runs,
            got.use_community_detection, got.max_shrink_factor) == (80, 2, 0, 3, 0, 3.0)
    # a bad value is an error and leaves the context unchanged
    p.write_text("c-t=40\nr-lp-rebalancing=maybe\n")
    L.mt_kahypar_configure_context_from_file(ctx, str(p).encode())
    assert b"maybe" in L.mkp_last_error()
    L.mt_kahypar_amd_get_params(ctx, C.byref(got))
    assert got.contraction_limit_multiplier == 80
    L.mt_kahypar_context_free(ctx)


def test_capi_errors_exit_like_the_reference(tmp_path):
    # default error mode: message on stderr and exit(-1) (macros.h:108)
    code = ("import ctypes as C, sys; sys.path.insert(0, %r); import mtk_amd; "
            "L = C.CDLL(mtk_amd.LIB_PATH); n = C.c_ulong(); p = C.c_void_p(); "
            "L.mt_kahypar_read_hypergraph_from_file(b'/nonexistent.hgr', C.byref(n), C.byref(n), C.byref(p), "
            "C.byref(p), C.byref(p), C.byref(p)); print('not reached')") % os.path.join(ROOT, "mt-kahypar_amd")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 255 and "not reached" not in r.stdout
    assert "[ERROR]" in r.stderr and "nonexistent" in r.stderr


def test_header_is_valid_c_and_cxx(tmp_path):
    for compiler, ext in (("gcc", "c"), ("g++", "cpp")):
        src = tmp_path / ("h." + ext)
        src.write_text('#include "libkahypar.h"\n#include "mtkahypar_amd.h"\nint main(void) { return 0; }\n')
        r = subprocess.run([compiler, "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"), "-c", str(src),
                            "-o", str(tmp_path / "h.o")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_request_estimate_grows_with_size_and_k():
    e = [mtk_amd.request_estimate(n, n, 6 * n, 2000, 8) for n in (1000, 10_000, 100_000)]
    assert e[0] < e[1] < e[2]
    assert mtk_amd.request_estimate(10_000, 10_000, 60_000, 2000, 128) >= e[1]
    assert e[2] >= 2 * 28 * 600_000  # at least the resident CSR, rounded to size classes
    with pytest.raises(RuntimeError):
        mtk_amd.request_estimate(10, 10, 60, 5, 1)
