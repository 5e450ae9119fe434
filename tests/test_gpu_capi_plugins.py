"""The drop-in boundaries on the GPU: the C library interface (libkahypar.h),
the C++ plugin interfaces (IRefiner / ICoarsener / run_parallel_louvain, via
a compiled program), and LP with explicit refinement nodes or without
rebalancing, each against the oracle or the one-call pipeline."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import mtk_amd
import oracle as orc
from helpers import GOLDEN, read_hgr, vectors

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROGRAMS = os.path.join(ROOT, "tests", "cpp", "_build")
IBM01 = os.path.join(GOLDEN, vectors()["lp_invariants"]["instance"])


@pytest.fixture(scope="module")
def ctx():
    c = mtk_amd.Context(0)
    yield c
    c.close()


def capi():
    L = C.CDLL(mtk_amd.LIB_PATH)
    L.mt_kahypar_context_new.restype = C.c_void_p
    L.mt_kahypar_context_free.argtypes = [C.c_void_p]
    L.mt_kahypar_configure_context_from_file.argtypes = [C.c_void_p, C.c_char_p]
    L.mt_kahypar_amd_set_error_mode.argtypes = [C.c_int]
    L.mt_kahypar_partition.argtypes = [C.c_ulong, C.c_ulong, C.c_double, C.c_uint, C.c_int, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.c_void_p, C.c_void_p,
                                       C.c_bool]
    L.mkp_last_error.restype = C.c_char_p
    return L


def capi_partition(L, ctx, n, noff, pins, vw, nw, k, eps, seed):
    idx = np.ascontiguousarray(noff, dtype=np.uint64)
    pn = np.ascontiguousarray(pins, dtype=np.uint64)
    part = np.zeros(n, np.uint32)
    obj = C.c_int()
    vwp = None if vw is None else np.ascontiguousarray(vw, dtype=np.int32).ctypes.data
    nwp = None if nw is None else np.ascontiguousarray(nw, dtype=np.int32).ctypes.data
    L.mt_kahypar_partition(n, len(idx) - 1, eps, k, seed, vwp, nwp, idx.ctypes.data, pn.ctypes.data,
                           C.byref(obj), ctx, part.ctypes.data, False)
    return part, obj.value


@pytest.mark.parametrize("k", [2, 4, 8])
def test_capi_partition_matches_oracle_pipeline(k):
    L = capi()
    L.mt_kahypar_amd_set_error_mode(1)
    n, noff, pins, vw, nw = read_hgr(IBM01)
    c = L.mt_kahypar_context_new()
    part, obj = capi_partition(L, c, n, noff, pins, vw, nw, k, 0.03, 42)
    L.mt_kahypar_context_free(c)
    assert obj >= 0, L.mkp_last_error()
    opart, okm1, _ = orc.Hypergraph(n, noff, pins, vw, nw).partition(k, 0.03, 42, sync=True)
    assert np.array_equal(part.astype(np.int32), opart)
    assert obj == okm1
    perfect = int(np.ceil(vw.sum() / k))
    assert np.bincount(part, weights=vw, minlength=k).max() <= int(1.03 * perfect)


def test_capi_ini_switches_lp_off(ctx, tmp_path):
    L = capi()
    L.mt_kahypar_amd_set_error_mode(1)
    n, noff, pins, vw, nw = read_hgr(IBM01)
    ini = tmp_path / "no_lp.ini"
    ini.write_text("r-lp-type=do_nothing\nc-t=100\n")
    c = L.mt_kahypar_context_new()
    L.mt_kahypar_configure_context_from_file(c, str(ini).encode())
    part, obj = capi_partition(L, c, n, noff, pins, None, None, 4, 0.03, 7)
    L.mt_kahypar_context_free(c)
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins)
    ref, st = dh.partition(4, 0.03, 7, lp_enabled=0, contraction_limit_multiplier=100)
    dh.close()
    assert np.array_equal(part.astype(np.int32), ref) and obj == st["km1"]
    assert st["lp_visits"] == 0


def test_capi_rejects_bad_input_with_return_mode():
    L = capi()
    L.mt_kahypar_amd_set_error_mode(1)
    c = L.mt_kahypar_context_new()
    part, obj = capi_partition(L, c, 3, [0, 2], [0, 5], None, None, 2, 0.03, 1)  # pin 5 >= 3 vertices
    assert obj == -1 and b"pin" in L.mkp_last_error()
    L.mt_kahypar_context_free(c)


def test_capi_program_in_c(tmp_path):
    exe = os.path.join(PROGRAMS, "capi_partition")
    out = tmp_path / "part.txt"
    r = subprocess.run([exe, IBM01, "4", "0.03", "42", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    n, noff, pins, vw, nw = read_hgr(IBM01)
    part = np.loadtxt(out, dtype=np.int32)
    opart, okm1, _ = orc.Hypergraph(n, noff, pins, vw, nw).partition(4, 0.03, 42, sync=True)
    assert np.array_equal(part, opart)
    assert ("objective=%d" % okm1) in r.stdout


@pytest.mark.parametrize("k,seed,fm", [(2, 1, False), (8, 5, False), (4, 3, True)])
def test_cxx_plugin_interfaces_match_pipeline(k, seed, fm):
    """LabelPropagationRefiner and MultiTryKWayFM through IRefiner, level by
    level, against the one-call pipeline."""
    exe = os.path.join(PROGRAMS, "plugin_pipeline")
    r = subprocess.run([exe, IBM01, str(k), str(seed)] + (["fm"] if fm else []), capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "SAME" in r.stdout
    if fm:
        assert "fm_improvement=0 " not in r.stdout


def test_capi_ini_switches_fm_on(ctx, tmp_path):
    L = capi()
    L.mt_kahypar_amd_set_error_mode(1)
    n, noff, pins, vw, nw = read_hgr(IBM01)
    ini = tmp_path / "fm.ini"
    ini.write_text("r-fm-type=fm_gain_cache\nr-fm-multitry-rounds=10\nr-fm-seed-nodes=25\n")
    c = L.mt_kahypar_context_new()
    L.mt_kahypar_configure_context_from_file(c, str(ini).encode())
    part, obj = capi_partition(L, c, n, noff, pins, vw, nw, 4, 0.03, 7)
    L.mt_kahypar_context_free(c)
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins, vw, nw)
    ref, st = dh.partition(4, 0.03, 7, fm_rounds=10)
    dh.close()
    assert np.array_equal(part.astype(np.int32), ref) and obj == st["km1"]
    assert st["fm_improvement"] > 0
    opart, okm1, ost = orc.Hypergraph(n, noff, pins, vw, nw).partition(4, 0.03, 7, sync=True, fm_rounds=10)
    assert np.array_equal(ref, opart) and okm1 == st["km1"] and ost["fm_improvement"] == st["fm_improvement"]


def random_state(n, noff, pins, k, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, k, n).astype(np.int32)


@pytest.mark.parametrize("k", [2, 8])
def test_lp_refinement_nodes_and_no_rebalancing(ctx, k):
    n, noff, pins, vw, nw = read_hgr(IBM01)
    part0 = random_state(n, noff, pins, k, 3)
    W = int(vw.sum())
    perfect = -(-W // k)
    maxw = [int(1.03 * perfect)] * k
    oh = orc.Hypergraph(n, noff, pins, vw, nw)
    dh = mtk_amd.Hypergraph(ctx, n, noff, pins, vw, nw)
    rng = np.random.default_rng(k)
    subset = rng.choice(n, size=n // 3, replace=True)  # duplicates are dropped on both sides
    for kw in (dict(nodes=subset), dict(rebalancing=False), dict(nodes=subset, rebalancing=False)):
        p = mtk_amd.Partition(dh, k)
        p.set_parts(part0)
        p.initialize()
        st = p.lp_refine(maxw, [perfect] * k, seed=9, **kw)
        opart, okm1, ost = oh.lp_refine(k, part0, maxw, [perfect] * k, seed=9, sync=True, **kw)
        assert np.array_equal(p.parts(), opart), kw
        assert p.km1() == okm1
        assert (st["moves"], st["reverts"], st["visits"], st["rounds"]) == \
            (ost["moves"], ost["reverts"], ost["visits"], ost["rounds"]), kw
        p.close()
    dh.close()
