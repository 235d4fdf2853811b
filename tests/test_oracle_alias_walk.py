"""Pinning the oracle's alias tables and walks against the closed forms the
reference's own tests use (CPU only)."""
import json
import os
from collections import defaultdict

import numpy as np
import pytest
from scipy import stats

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- alias ----
def test_alias_hand_worked_cases():
    cases = json.load(open(os.path.join(GOLD, "alias_hand.json")))["cases"]
    for c in cases:
        w = np.array(c["w"], np.float32)
        d = len(w)
        rp = np.array([0, d], np.uint64)
        col = np.arange(100, 100 + d, dtype=np.uint32)   # neighbour id = 100 + edge index
        p, prim, al, meta, vs = O.alias_build(rp, col, w)
        assert vs[0] == np.float32(c["sum"])
        for k, (bp, pi, qi) in enumerate(c["buckets"]):
            assert p[k] == np.float32(bp), (c["w"], k)
            assert prim[k] == 100 + pi
            assert al[k] == (100 + qi if qi >= 0 else 100 + pi)


def test_alias_mass_conservation_rmat():
    rp, col, w = O.rmat_graph(10, 16, seed=1, weighted=True)
    p, prim, al, meta, vs = O.alias_build(rp, col, w)
    assert np.all((p > 0) & (p <= 1))
    V = len(rp) - 1
    for v in range(V):
        b, e = int(rp[v]), int(rp[v + 1])
        d = e - b
        if d == 0:
            continue
        mass = defaultdict(float)
        for k in range(b, e):
            mass[int(prim[k])] += float(p[k])
            if al[k] != prim[k] or p[k] < 1:
                mass[int(al[k])] += 1.0 - float(p[k])
        tot = float(np.sum(w[b:e], dtype=np.float64))
        # The reference's float32 running sum and repeated subtraction carry
        # rounding error that grows with the degree: allow 8 d ulps (a lost or
        # misplaced piece is off by a sizeable fraction of a bucket instead).
        tol = 8 * d * 2.0 ** -24
        assert abs(float(vs[v]) - tot) <= tol * tot
        for k in range(b, e):
            want = d * float(w[k]) / float(vs[v])
            assert abs(mass[int(col[k])] - want) <= tol * max(1.0, want), (v, k)


# ------------------------------------------------------- helpers ----
def chi2_rows(counts_by_row, probs_by_row, min_expected=5.0):
    """Pooled chi-square over rows: sum of per-row statistics on cells with
    enough expected mass; returns (statistic, dof, p-value)."""
    stat, dof = 0.0, 0
    for key, cnt in counts_by_row.items():
        probs = probs_by_row[key]
        n = sum(cnt.values())
        if n == 0:
            continue
        for x in cnt:
            assert probs.get(x, 0.0) > 0.0, f"transition {key}->{x} has zero probability"
        cells = [(x, pr * n) for x, pr in probs.items()]
        big = [(x, e) for x, e in cells if e >= min_expected]
        small_e = sum(e for x, e in cells if e < min_expected)
        small_o = sum(cnt.get(x, 0) for x, e in cells if e < min_expected)
        for x, e in big:
            stat += (cnt.get(x, 0) - e) ** 2 / e
        k = len(big)
        if small_e >= min_expected:
            stat += (small_o - small_e) ** 2 / small_e
            k += 1
        dof += max(k - 1, 0)
    return stat, dof, stats.chi2.sf(stat, dof) if dof else 1.0


def undirected_random_graph(V, E, seed, weighted=True):
    """Like the reference test generator (src/tests/test.hpp:49-99): random
    undirected pairs, deduplicated, mirrored with the same data; weights
    rand()%5+1 (generator_helper.hpp:65-76)."""
    rng = np.random.default_rng(seed)
    pairs = set()
    while len(pairs) < E:
        a, b = rng.integers(0, V, 2)
        if a != b:
            pairs.add((min(a, b), max(a, b)))
    src, dst, w = [], [], []
    for a, b in sorted(pairs):
        x = float(rng.integers(1, 6))
        src += [a, b]
        dst += [b, a]
        w += [x, x]
    return O.csr_from_edges(V, np.array(src), np.array(dst), np.array(w, np.float32) if weighted else None,
                            False, True)


def static_probs(rp, col, w):
    P = {}
    for v in range(len(rp) - 1):
        b, e = int(rp[v]), int(rp[v + 1])
        if e == b:
            continue
        ws = w[b:e].astype(np.float64) if w is not None else np.ones(e - b)
        P[v] = defaultdict(float)
        for k in range(b, e):
            P[v][int(col[k])] += ws[k - b] / ws.sum()
    return P


def transitions(paths, lens, order=1):
    cnt = defaultdict(lambda: defaultdict(int))
    for row, n in zip(paths, lens):
        r = row[:n].tolist()
        for i in range(len(r) - order):
            key = r[i] if order == 1 else (r[i], r[i + 1])
            cnt[key][r[i + order]] += 1
    return cnt


# --------------------------------------------------------------- walks ----
@pytest.mark.parametrize("weighted", [True, False])
def test_deepwalk_static_distribution(karate, weighted):
    """P(v -> u) = w(v,u) / sum w (src/tests/test_walk.hpp:130-186)."""
    src, dst, w = karate
    rp, col, ww = O.csr_from_edges(34, src, dst, w, False, False)
    g = O.Graph(rp, col, ww)
    paths, lens, tr, st = O.walk(g, O.DEEPWALK, 34 * 300, 40, weighted, seed=3)
    assert np.all(lens == 41) and not tr.any()
    stat, dof, pv = chi2_rows(transitions(paths, lens), static_probs(rp, col, ww if weighted else None))
    assert pv > 1e-4, (stat, dof)


def node2vec_probs(rp, col, w, p, q):
    """Closed form of src/tests/test_node2vec.cpp:45-79 per directed edge (t -> v):
    P(x) proportional to w(v,x) * (1/p if x == t, 1 if x in N(t), 1/q otherwise)."""
    V = len(rp) - 1
    nbrs = [set(col[rp[v]:rp[v + 1]].tolist()) for v in range(V)]
    P = {}
    for t in range(V):
        for v in nbrs[t]:
            b, e = int(rp[v]), int(rp[v + 1])
            row = defaultdict(float)
            for k in range(b, e):
                x = int(col[k])
                f = 1 / p if x == t else (1.0 if x in nbrs[t] else 1 / q)
                row[x] += f * (float(w[k]) if w is not None else 1.0)
            s = sum(row.values())
            P[(t, v)] = {x: y / s for x, y in row.items()}
    return P


@pytest.mark.parametrize("p,q,weighted", [(2.0, 0.5, True), (2.0, 0.5, False), (0.25, 2.0, True),
                                          (0.5, 4.0, False), (1.0, 1.0, True), (4.0, 0.25, True)])
def test_node2vec_second_order_distribution(p, q, weighted):
    rp, col, w = undirected_random_graph(40, 160, seed=int(p * 8 + q * 2), weighted=True)
    g = O.Graph(rp, col, w)
    paths, lens, tr, st = O.walk(g, O.NODE2VEC, 40 * 60, 80, weighted, p=p, q=q, seed=5)
    probs = node2vec_probs(rp, col, w if weighted else None, p, q)
    stat, dof, pv = chi2_rows(transitions(paths, lens, order=2), probs)
    assert pv > 1e-4, (stat, dof, pv)


def test_closed_form_catches_a_planted_mistake():
    """The node2vec check has teeth: walks made with p and q exchanged (the kind
    of slip a port could make) fail it."""
    rp, col, w = undirected_random_graph(40, 160, seed=77)
    g = O.Graph(rp, col, w)
    paths, lens, tr, st = O.walk(g, O.NODE2VEC, 40 * 60, 80, True, p=0.5, q=2.0, seed=5)
    _, _, pv = chi2_rows(transitions(paths, lens, order=2), node2vec_probs(rp, col, w, 2.0, 0.5))
    assert pv < 1e-12


def test_node2vec_trials_per_step_matches_closed_form():
    """Expected trials per step (paper p.6 Eq. 3): for a step from edge (t,v) the
    acceptance probability is sum_x w(v,x) P_d(x) / (upper * sum_x w(v,x))."""
    rp, col, w = undirected_random_graph(30, 120, seed=9)
    g = O.Graph(rp, col, w)
    p, q = 2.0, 0.5
    n = 30 * 200
    paths, lens, tr, st = O.walk(g, O.NODE2VEC, n, 40, True, p=p, q=q, seed=8)
    V = len(rp) - 1
    nbrs = [set(col[rp[v]:rp[v + 1]].tolist()) for v in range(V)]
    upper = max(1.0, 1 / q)
    exp_trials = 0.0
    for row, m in zip(paths, lens):
        r = row[:m].tolist()
        exp_trials += 1.0   # step 0 always accepts
        for i in range(1, len(r) - 1):
            t, v = r[i - 1], r[i]
            b, e = int(rp[v]), int(rp[v + 1])
            num = sum(float(w[k]) * (1 / p if col[k] == t else (1.0 if col[k] in nbrs[t] else 1 / q))
                      for k in range(b, e))
            den = upper * float(np.sum(w[b:e], dtype=np.float64))
            exp_trials += den / num
    assert abs(st.trials - exp_trials) < 5 * np.sqrt(exp_trials)


def test_ppr_geometric_lengths(karate):
    """Continue with probability 1-t at each step (ppr.hpp:36, walk.hpp:963-968):
    steps ~ Geometric, P(k) = (1-t)^k t, on a graph without isolated vertices."""
    src, dst, w = karate
    rp, col, ww = O.csr_from_edges(34, src, dst, w, False, False)
    g = O.Graph(rp, col, ww)
    t = 0.15
    paths, lens, tr, st = O.walk(g, O.PPR, 20000, 0, True, terminate_prob=t, seed=4)
    steps = lens - 1
    K = 30
    obs = np.bincount(np.minimum(steps, K), minlength=K + 1)
    pk = np.array([(1 - t) ** k * t for k in range(K)] + [(1 - t) ** K])
    assert stats.chisquare(obs, pk * len(steps)).pvalue > 1e-4
    assert not tr.any()
    # and the path is a static walk
    stat, dof, pv = chi2_rows(transitions(paths, lens), static_probs(rp, col, ww))
    assert pv > 1e-4


def test_ppr_start_list_and_cap():
    rp, col, w = undirected_random_graph(20, 60, seed=2)
    g = O.Graph(rp, col, w)
    starts = np.array([3, 7], np.uint32)
    paths, lens, tr, st = O.walk(g, O.PPR, 100, 5, True, terminate_prob=0.0, seed=1, start_mode=O.START_LIST,
                                 starts=starts)
    assert np.array_equal(paths[:, 0], starts[np.arange(100) % 2])
    deg = np.diff(rp.astype(np.int64))
    nonisolated = deg[paths[:, 0]] > 0
    assert np.all(lens[nonisolated] == 6) and np.all(tr[nonisolated])


def test_isolated_start_gives_one_vertex_path():
    """A walker starting at an isolated vertex yields [start] (test_bound.cpp:223)."""
    src = np.array([0, 1], np.uint32)
    dst = np.array([1, 0], np.uint32)
    rp, col, w = O.csr_from_edges(3, src, dst, np.ones(2, np.float32), False, True)
    g = O.Graph(rp, col, w)
    for kind in (O.DEEPWALK, O.NODE2VEC):
        paths, lens, tr, st = O.walk(g, kind, 3, 10, True, p=2.0, q=0.5, seed=1)
        assert list(lens) == [11, 11, 1] and paths[2, 0] == 2
        assert list(paths[0, :4]) == [0, 1, 0, 1]


def test_epochs_give_the_same_rows():
    """Splitting walkers into epochs (WalkConfig::rate, walk.hpp:706-712) does
    not change any path: draws are keyed by walker id."""
    rp, col, w = undirected_random_graph(50, 200, seed=3)
    g = O.Graph(rp, col, w)
    full, lf, _, _ = O.walk(g, O.NODE2VEC, 300, 20, True, p=2.0, q=0.5, seed=6)
    a, la, _, _ = O.walk(g, O.NODE2VEC, 120, 20, True, p=2.0, q=0.5, seed=6)
    b, lb, _, _ = O.walk(g, O.NODE2VEC, 180, 20, True, p=2.0, q=0.5, seed=6, walker_begin=120)
    assert np.array_equal(np.concatenate([la, lb]), lf)
    assert np.array_equal(np.concatenate([a, b]), full)


def metapath_graph(seed=11):
    rp, col, w = undirected_random_graph(60, 300, seed=seed)
    vt = np.random.default_rng(seed).integers(0, 5, 60).astype(np.uint8)
    et = vt[col]
    return rp, col, w, et


SCHEMES = [[0b00011, 0b11100], [0b10101, 0b01010, 0b11111]]


@pytest.mark.parametrize("weighted", [True, False])
def test_metapath_distribution(weighted):
    """Every step takes an edge whose type the scheme state allows, with
    probability proportional to the weight among allowed edges; the walk stops
    at L or where no out-edge type is allowed (metapath.hpp:101-121,
    src/tests/test_metapath.cpp:131-176)."""
    rp, col, w, et = metapath_graph()
    g = O.Graph(rp, col, w, et)
    n = 60 * 100
    paths, lens, tr, st = O.walk(g, O.METAPATH, n, 20, weighted, seed=12, schemes=SCHEMES)
    # recover each walker's scheme from the draw it makes at init (tag 2, step 0)
    cnt = defaultdict(lambda: defaultdict(int))
    probs = {}
    vmask = O.vmask_of(rp, et)
    for i in range(n):
        r = O.philox([i, 0, 0, 2], [12, 0])
        s = O.bounded(int(r[0]), int(r[1]), len(SCHEMES))
        sch = SCHEMES[s]
        path = paths[i, :lens[i]].tolist()
        state = 0
        for k in range(len(path) - 1):
            v, x = path[k], path[k + 1]
            m = sch[state]
            b, e = int(rp[v]), int(rp[v + 1])
            key = (v, m)
            if key not in probs:
                allowed = [(int(col[j]), float(w[j]) if weighted else 1.0) for j in range(b, e)
                           if (m >> int(et[j])) & 1]
                tot = sum(y for _, y in allowed)
                probs[key] = defaultdict(float)
                for u, y in allowed:
                    probs[key][u] += y / tot
            cnt[key][x] += 1
            state = (state + 1) % len(sch)
        # termination rule
        v = path[-1]
        if lens[i] - 1 < 20:
            assert not (vmask[v] & sch[state]) or rp[v + 1] == rp[v]
    stat, dof, pv = chi2_rows(cnt, probs)
    assert pv > 1e-4, (stat, dof)


def test_uniform_start_distribution():
    rp, col, w = undirected_random_graph(16, 40, seed=4)
    g = O.Graph(rp, col, w)
    paths, lens, tr, st = O.walk(g, O.DEEPWALK, 16000, 1, True, seed=2, start_mode=O.START_UNIFORM)
    obs = np.bincount(paths[:, 0], minlength=16)
    assert stats.chisquare(obs).pvalue > 1e-4
