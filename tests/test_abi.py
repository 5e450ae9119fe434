"""C-ABI library checks that need no GPU: the library loads, exports every
function the headers declare, and the host-side generators behave."""
import ctypes
import glob
import os
import re

import numpy as np

import mtk_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"(?:MKP_API|KAHYPAR_API)\s+[\w\s\*]+?\b((?:mkp|mt_kahypar)_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(mtk_amd.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_binding_signatures_cover_header():
    names = {n for n in declared_symbols() if n.startswith("mkp_")}
    assert names <= set(mtk_amd.SIGNATURES), sorted(names - set(mtk_amd.SIGNATURES))


def test_powerlaw_generator_properties():
    h = mtk_amd.HostHypergraph.powerlaw(4096, 4096, 2.5, 2, 64, window=1024, local_fraction=0.8, seed=42)
    a = h.arrays()
    sizes = np.diff(a["noff"].astype(np.int64))
    assert sizes.min() >= 2 and sizes.max() <= 64
    for e in range(0, 4096, 97):
        net = a["pins"][a["noff"][e]:a["noff"][e + 1]]
        assert len(set(net.tolist())) == len(net)
    assert a["pins"].max() < 4096
    b = mtk_amd.HostHypergraph.powerlaw(4096, 4096, 2.5, 2, 64, window=1024, local_fraction=0.8, seed=42).arrays()
    assert np.array_equal(a["pins"], b["pins"])  # seeded and deterministic
    c = mtk_amd.HostHypergraph.powerlaw(4096, 4096, 2.5, 2, 64, window=1024, local_fraction=0.8, seed=43).arrays()
    assert not np.array_equal(a["noff"], c["noff"])
    # mean of the truncated power law 2.5 on [2, 64]
    s = np.arange(2, 65)
    pmf = s ** -2.5 / (s ** -2.5).sum()
    assert abs(sizes.mean() - (pmf * s).sum()) < 0.2


def test_rownet_generator_properties():
    h = mtk_amd.HostHypergraph.rownet(20000, 20000, 8.0, 500, 2000, 0.7, seed=1)
    a = h.arrays()
    sizes = np.diff(a["noff"].astype(np.int64))
    assert sizes.min() >= 2  # empty and single-pin columns dropped
    assert abs(a["vw"].mean() - 8.0) < 0.6  # vertex weight = row nnz
    assert a["vw"].sum() >= len(a["pins"])


def test_errors_are_reported_not_fatal():
    h = ctypes.c_void_p()
    prm = mtk_amd.PowerlawParams(1, 1, 2.0, 2, 1, 0, 0.0, 1, 1, 0)
    st = mtk_amd.lib().mkp_generate_powerlaw(ctypes.byref(prm), ctypes.byref(h))
    assert st == 1
    assert b"bad sizes" in mtk_amd.lib().mkp_last_error()


def test_partition_params_out_of_range_are_refused_before_any_work():
    """Sub-round counts above 255 would overflow the 8-bit sub-round field
    of the round tags and the 256-entry bucket tables: the pipeline and the
    server refuse them with an invalid-argument status before touching the
    hypergraph (the handles below are never dereferenced)."""
    L = mtk_amd.lib()
    dummy = ctypes.create_string_buffer(64)
    hg = ctypes.cast(dummy, ctypes.c_void_p)
    parts = np.zeros(4, np.int32)
    for field in ("lp_sub_rounds", "louvain_sub_rounds", "coarsening_sub_rounds"):
        for bad in (256, 1000, -1):
            prm = mtk_amd.partition_params(4, 0.03, 1, **{field: bad})
            st = L.mkp_partition_hypergraph(hg, ctypes.byref(prm), parts.ctypes.data_as(ctypes.c_void_p), None)
            assert st == 1, (field, bad)
            assert field.encode() in L.mkp_last_error()
            t = ctypes.c_uint64()
            st = L.mkp_server_submit(hg, hg, ctypes.byref(prm), ctypes.byref(t))
            assert st == 1, (field, bad)
            assert field.encode() in L.mkp_last_error()
