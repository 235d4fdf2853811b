"""Pinning the oracle's random numbers and graph builders (CPU only)."""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def kat_vectors():
    rows = []
    for line in open(os.path.join(GOLD, "philox_kat.txt")):
        if line.startswith("#") or not line.strip():
            continue
        f = line.split()
        assert f[0] == "philox4x32" and f[1] == "10"
        vals = [int(x, 16) for x in f[2:]]
        rows.append((vals[0:4], vals[4:6], vals[6:10]))
    return rows


@pytest.mark.parametrize("ctr,key,want", kat_vectors())
def test_philox_known_answers(ctr, key, want):
    assert list(O.philox(ctr, key)) == want


def test_u01_grid():
    assert O.u01(0) == 0.0
    assert O.u01(0xFFFFFFFF) == np.float32(1 - 2.0 ** -24)
    rng = np.random.default_rng(1)
    for u in rng.integers(0, 2 ** 32, 2000, dtype=np.uint64):
        assert Fraction(O.u01(int(u))) == Fraction(int(u) >> 8, 2 ** 24)


def test_bounded_is_floor_of_product():
    rng = np.random.default_rng(2)
    cases = [(0, 0, 1), (0xFFFFFFFF, 0xFFFFFFFF, 1), (0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF),
             (0, 0x80000000, 3), (0xFFFFFFFF, 0x7FFFFFFF, 2)]
    cases += [(int(a), int(b), int(n)) for a, b, n in zip(rng.integers(0, 2 ** 32, 3000),
                                                          rng.integers(0, 2 ** 32, 3000),
                                                          rng.integers(1, 2 ** 32, 3000))]
    for lo, hi, n in cases:
        r = (hi << 32) | lo
        assert O.bounded(lo, hi, n) == (r * n) >> 64
        assert O.bounded(lo, hi, n) < n


def test_rmat_bit_statistics():
    """Each level picks quadrant a/b/c/d = 0.57/0.19/0.19/0.05: the src bit is set
    with probability c+d, the dst bit with b+d, both with d."""
    s = 12
    src, dst = O.rmat_raw(s, 16, seed=3)
    n = len(src)
    for lvl in range(s):
        bit = 1 << (s - 1 - lvl)
        sb = (src & bit) != 0
        db = (dst & bit) != 0
        for got, p in ((sb.mean(), 0.24), (db.mean(), 0.24), ((sb & db).mean(), 0.05)):
            assert abs(got - p) < 5 * np.sqrt(p * (1 - p) / n)
    assert src.max() < 2 ** s and dst.max() < 2 ** s


def test_rmat_seed_changes_graph():
    a = O.rmat_raw(8, 4, seed=1)
    b = O.rmat_raw(8, 4, seed=2)
    assert not np.array_equal(a[0], b[0])


@pytest.mark.parametrize("scale", [6, 9])
def test_build_undirected_brute_force(scale):
    V = 1 << scale
    src, dst = O.rmat_raw(scale, 16, seed=5)
    rp, col, w = O.build_undirected(V, src, dst, seed=5, weighted=True)
    want = set()
    for a, b in zip(src.tolist(), dst.tolist()):
        if a != b:
            want.add((a, b))
            want.add((b, a))
    got = []
    for v in range(V):
        lst = col[rp[v]:rp[v + 1]]
        assert np.all(np.diff(lst.astype(np.int64)) > 0), "lists sorted, no duplicates"
        got += [(v, int(u)) for u in lst]
    assert set(got) == want and len(got) == len(want)
    # weights: per undirected edge, equal in both directions, in [1,5)
    wmap = {}
    for v in range(V):
        for k in range(rp[v], rp[v + 1]):
            wmap[(v, int(col[k]))] = w[k]
    for (a, b), x in wmap.items():
        assert x == wmap[(b, a)] and 1.0 <= x < 5.0
        assert x == O.edge_weight(5, a, b)


def _csr_reference_semantics(V, src, dst, w, undirected, sort_adj):
    """graph.hpp:430-443 (mirror: e, reversed e, in file order), 309-330 (CSR in
    file order per source), node2vec.hpp:122-125 (sort by neighbour)."""
    edges = []
    for i in range(len(src)):
        edges.append((int(src[i]), int(dst[i]), float(w[i])))
        if undirected:
            edges.append((int(dst[i]), int(src[i]), float(w[i])))
    lists = [[] for _ in range(V)]
    for s, d, x in edges:
        lists[s].append((d, x))
    if sort_adj:
        lists = [sorted(l, key=lambda t: t[0]) for l in lists]
    rp = np.cumsum([0] + [len(l) for l in lists]).astype(np.uint64)
    col = np.array([d for l in lists for d, _ in l], np.uint32)
    ww = np.array([x for l in lists for _, x in l], np.float32)
    return rp, col, ww


@pytest.mark.parametrize("undirected,sort_adj", [(False, False), (True, False), (True, True), (False, True)])
def test_csr_from_edges_karate(karate, undirected, sort_adj):
    src, dst, w = karate
    if undirected:
        keep = src < dst
        src, dst, w = src[keep], dst[keep], w[keep]
    # shuffle the file order so that order preservation is visible
    perm = np.random.default_rng(4).permutation(len(src))
    src, dst, w = src[perm], dst[perm], w[perm]
    got = O.csr_from_edges(34, src, dst, w, undirected, sort_adj)
    want = _csr_reference_semantics(34, src, dst, w, undirected, sort_adj)
    for g, x in zip(got, want):
        assert np.array_equal(g, x)


def test_vertex_types_uniform():
    vt = O.vertex_types(50000, 5, 9)
    assert vt.max() == 4
    counts = np.bincount(vt, minlength=5)
    assert np.all(np.abs(counts - 10000) < 5 * np.sqrt(10000))


def test_node2vec_constants_follow_reference_rounding():
    # upper = max(1, 1/q), lower = min(1/p, 1, 1/q) in double, stored as float
    up, lo, ip, iq, ov = O.node2vec_constants(2.0, 0.5)
    assert (up, lo, ip, iq) == (2.0, 0.5, 0.5, 2.0) and ov == -1.5
    up, lo, ip, iq, ov = O.node2vec_constants(0.25, 2.0)
    assert up == 1.0 and lo == 0.5 and ip == 4.0 and ov == 3.0
    up, lo, ip, iq, ov = O.node2vec_constants(3.0, 3.0)
    assert up == 1.0 and lo == np.float32(1 / 3.0) and iq == np.float32(1) / np.float32(3)


def test_package_node2vec_constants_equal_the_oracle():
    """The host constants knightking_amd hands to walk programs are the
    oracle's, bit for bit, over a grid of p and q (including values whose
    reciprocals round differently in float and double)."""
    import knightking_amd as kk
    vals = [0.1, 0.25, 0.3, 0.5, 0.7, 1.0, 1.5, 2.0, 3.0, 7.0, 1e-3, 123.456]
    for p in vals:
        for q in vals:
            got = np.array(kk.node2vec_constants(p, q), np.float32)
            want = np.array(O.node2vec_constants(p, q), np.float32)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (p, q, got, want)


def test_uniform_start_is_the_kernels_draw():
    """kkw_uniform_start (get_uniform_dist_func on the host) is the start draw
    of the kernels: Philox at (walker, 0, 0, INIT), words z, w mapped to
    [0, v_num) -- checked against the oracle's Philox and bounded map."""
    import knightking_amd as kk
    for seed in (1, 7, (3 << 40) + 5):
        for v_num in (1, 34, 1000, 1 << 24, 0xFFFFFFFF):
            for wid in (0, 1, 33, 12345, 0xFFFFFFFE):
                r = O.philox([wid, 0, 0, 2], [seed & 0xFFFFFFFF, seed >> 32])
                assert kk.uniform_start(seed, wid, v_num) == O.bounded(int(r[2]), int(r[3]), v_num)
    assert kk.uniform_start(1, 5, 0) == 0
