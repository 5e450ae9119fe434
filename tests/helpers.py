"""Shared test helpers: fixture loading and small hypergraph constructors."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def vectors():
    with open(os.path.join(GOLDEN, "reference_vectors.json")) as f:
        return json.load(f)


def csr(nets):
    noff = np.zeros(len(nets) + 1, np.uint64)
    for i, e in enumerate(nets):
        noff[i + 1] = noff[i] + len(e)
    pins = np.array([p for e in nets for p in e], dtype=np.uint32)
    return noff, pins


def read_hgr(path):
    """hMetis reader following io/hypergraph_io.cpp:95-177: 1-based pins,
    header 'E N [fmt]' (1 = net weights, 10 = vertex weights, 11 = both),
    single-pin nets dropped."""
    with open(path) as f:
        lines = [l for l in (x.strip() for x in f) if l and not l.startswith("%")]
    head = lines[0].split()
    m, n = int(head[0]), int(head[1])
    fmt = int(head[2]) if len(head) > 2 else 0
    nets, nw = [], []
    for i in range(m):
        vals = [int(x) for x in lines[1 + i].split()]
        w = 1
        if fmt in (1, 11):
            w, vals = vals[0], vals[1:]
        if len(vals) > 1:
            nets.append([v - 1 for v in vals])
            nw.append(w)
    vw = np.ones(n, np.int32)
    if fmt in (10, 11):
        vw = np.array([int(lines[1 + m + i]) for i in range(n)], dtype=np.int32)
    noff, pins = csr(nets)
    return n, noff, pins, vw, np.array(nw, dtype=np.int32)
