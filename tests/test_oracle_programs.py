"""Walk programs on the oracle (oracle/program_oracle.cpp, compiled with each
program's source).  The generic loop is pinned two ways:

- programs that restate the built-in apps (knightking_amd/programs/deepwalk.cpp,
  ppr.cpp, node2vec.cpp -- node2vec with its outlier appendix) give the
  built-in oracle walks bit for bit: two independent codings of the same
  counters and the same float arithmetic;
- the tag walks of the reference's bound and outlier tests
  (src/tests/test_bound.cpp, test_outlier.cpp) and the return-biased walk of
  src/examples/dynamic_walk.cpp follow their closed-form transition laws
  (pooled chi-square), and a planted mistake in a program is caught.
"""
import os
from collections import defaultdict

import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_alias_walk import chi2_rows, undirected_random_graph

_TESTS = os.path.dirname(os.path.abspath(__file__))
PROGRAM_DIRS = [os.path.join(os.path.dirname(_TESTS), "knightking_amd", "programs"), os.path.join(_TESTS, "programs")]


def program_path(name):
    """Example programs ship in knightking_amd/programs; the reference-test
    programs (tag walks) live beside the tests."""
    for d in PROGRAM_DIRS:
        p = os.path.join(d, name + ".cpp")
        if os.path.exists(p):
            return p
    raise FileNotFoundError(name)
TAGS = 4


@pytest.fixture(scope="module")
def build_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("program_oracles"))


def plib(name, build_dir):
    return O.program_oracle(program_path(name), build_dir)


def same_paths(pa, la, pb, lb):
    assert np.array_equal(la, lb)
    n = pa.shape[1] if pa.shape[1] < pb.shape[1] else pb.shape[1]
    mask = np.arange(n)[None, :] < la[:, None]
    assert np.all((pa[:, :n] == pb[:, :n]) | ~mask)


# ------------------------------------------------ programs = built-ins ----
@pytest.mark.parametrize("weighted", [True, False])
def test_deepwalk_program_is_the_builtin(build_dir, weighted):
    rp, col, w = O.rmat_graph(10, 16, seed=3, weighted=True)
    L = 30
    g = O.Graph(rp, col, w)
    pb, lb, tb, _ = O.walk(g, O.DEEPWALK, 5000, L, weighted, seed=9)
    pa, la, ta = O.program_walk(plib("deepwalk", build_dir), rp, col, w, params=[L], use_weights=weighted,
                                seed=9, walker_num=5000, max_length=L)
    same_paths(pa, la, pb, lb)
    assert np.array_equal(ta, tb)


@pytest.mark.parametrize("t", [0.15, 1.0 / 80])
def test_ppr_program_is_the_builtin(build_dir, t):
    rp, col, w = O.rmat_graph(10, 16, seed=4, weighted=True)
    cap = O.ppr_default_cap(t)
    g = O.Graph(rp, col, w)
    starts = np.array([3, 17, 40], np.uint32)
    pb, lb, tb, _ = O.walk(g, O.PPR, 4000, cap, True, terminate_prob=t, seed=2, start_mode=O.START_LIST,
                           starts=starts)
    ext = np.float32(1.0) - np.float32(t)
    pa, la, ta = O.program_walk(plib("ppr", build_dir), rp, col, w, params=[ext], use_weights=True, seed=2,
                                walker_num=4000, max_length=cap, start_vertices=starts)
    same_paths(pa, la, pb, lb)
    assert np.array_equal(ta, tb)


@pytest.mark.parametrize("p,q,weighted", [(2.0, 0.5, True), (0.5, 2.0, True), (0.25, 1.0, False)])
def test_node2vec_program_is_the_builtin(build_dir, p, q, weighted):
    """(0.5, 2) and (0.25, 1) put 1/p above the upper bound: the outlier
    appendix of both codings is exercised."""
    rp, col, w = O.rmat_graph(10, 16, seed=5, weighted=True)
    L = 40
    g = O.Graph(rp, col, w)
    pb, lb, tb, _ = O.walk(g, O.NODE2VEC, 3000, L, weighted, p=p, q=q, seed=11)
    up, lo, ip, iq, ov = O.node2vec_constants(p, q)
    pa, la, ta = O.program_walk(plib("node2vec", build_dir), rp, col, w,
                                params=[L, up, lo, ip, iq, ov, 1.0 if weighted else 0.0], use_weights=weighted,
                                seed=11, walker_num=3000, max_length=L, second_order=True)
    same_paths(pa, la, pb, lb)
    assert np.array_equal(ta, tb)


@pytest.mark.parametrize("p,q,weighted", [(2.0, 0.5, True), (0.5, 2.0, True), (0.25, 1.0, False)])
def test_node2vec_query_program_is_the_builtin(build_dir, p, q, weighted):
    """node2vec written with the second-order query / response functions
    (knightking_amd/programs/node2vec_query.cpp: the query goes to the
    previous vertex, which answers by adjacency) gives the built-in walk bit
    for bit, outlier appendix included."""
    rp, col, w = O.rmat_graph(10, 16, seed=5, weighted=True)
    L = 40
    g = O.Graph(rp, col, w)
    pb, lb, tb, _ = O.walk(g, O.NODE2VEC, 3000, L, weighted, p=p, q=q, seed=11)
    up, lo, ip, iq, ov = O.node2vec_constants(p, q)
    pa, la, ta = O.program_walk(plib("node2vec_query", build_dir), rp, col, w,
                                params=[L, up, lo, ip, iq, ov, 1.0 if weighted else 0.0], use_weights=weighted,
                                seed=11, walker_num=3000, max_length=L)
    same_paths(pa, la, pb, lb)
    assert np.array_equal(ta, tb)


def test_query_sent_to_the_wrong_vertex_is_caught(build_dir, tmp_path):
    """A planted mistake in the query's target (the candidate itself instead
    of the previous vertex) changes the paths: the comparison above can fail."""
    src = open(program_path("node2vec_query")).read()
    assert "target = s.prev;" in src
    bad = tmp_path / "node2vec_query_bad.cpp"
    bad.write_text(src.replace("target = s.prev;", "target = e.neighbour;", 1))
    rp, col, w = O.rmat_graph(10, 16, seed=5, weighted=True)
    up, lo, ip, iq, ov = O.node2vec_constants(2.0, 0.5)
    g = O.Graph(rp, col, w)
    pb, lb, _, _ = O.walk(g, O.NODE2VEC, 3000, 40, True, p=2.0, q=0.5, seed=11)
    pa, la, _ = O.program_walk(O.program_oracle(str(bad), build_dir), rp, col, w,
                               params=[40, up, lo, ip, iq, ov, 1.0], use_weights=True, seed=11, walker_num=3000,
                               max_length=40)
    mask = np.arange(41)[None, :] < lb[:, None]
    assert not (np.array_equal(la, lb) and np.all((pa[:, :41] == pb) | ~mask))


# ----------------------------------------------------- closed forms ----
def tagged_graph(V, E, seed, weighted, sort_by_tag=False):
    """Random undirected graph, one random tag per CSR entry (both directions
    drawn independently, as the reference's generator does per edge record),
    random vertex tags; optionally each list ordered by (tag, neighbour)."""
    rp, col, w = undirected_random_graph(V, E, seed, weighted)
    rng = np.random.default_rng(seed + 100)
    et = rng.integers(0, TAGS, len(col)).astype(np.uint8)
    vt = rng.integers(0, TAGS, V).astype(np.int32)
    if sort_by_tag:
        for v in range(V):
            b, e = int(rp[v]), int(rp[v + 1])
            order = np.lexsort((col[b:e], et[b:e]))
            col[b:e], et[b:e] = col[b:e][order], et[b:e][order]
            if w is not None:
                w[b:e] = w[b:e][order]
    return rp, col, w, et, vt


def tag_transitions(paths, lens, vt):
    """Counts keyed by (cur, walker tag, previous vertex's tag) -> next vertex,
    steps >= 1 (the state is defined once a previous vertex exists); walker
    tag at step s = (id + s) % TAGS."""
    cnt = defaultdict(lambda: defaultdict(int))
    for i, (row, n) in enumerate(zip(paths, lens)):
        r = row[:n].tolist()
        for s in range(1, len(r) - 1):
            cnt[(r[s], (i + s) % TAGS, int(vt[r[s - 1]]))][r[s + 1]] += 1
    return cnt


def tag_probs(rp, col, w, et, vt, keys, weight_of_state):
    P = {}
    for key in keys:
        v, wt, pt = key
        b, e = int(rp[v]), int(rp[v + 1])
        vals = {}
        for k in range(b, e):
            sw = float(w[k]) if w is not None else 1.0
            vals[int(col[k])] = vals.get(int(col[k]), 0.0) + sw * weight_of_state(v, wt, pt, int(et[k]))
        tot = sum(vals.values())
        P[key] = {x: y / tot for x, y in vals.items()}
    return P


@pytest.mark.parametrize("weighted", [True, False])
def test_tag_walk_follows_its_closed_form(build_dir, weighted):
    """src/tests/test_bound.cpp:195-374: P(e | v, walker tag, previous tag) is
    proportional to static(e) * (1 + (prev tag + walker tag + edge tag) % 4
    + tag(v) * vertex_tag_weight), with lower and upper bounds that depend on v."""
    V, E = 40, 140
    rp, col, w, et, vt = tagged_graph(V, E, 21, weighted)
    vw = 0.5
    L = 20
    pa, la, _ = O.program_walk(plib("tag_walk", build_dir), rp, col, w, et, arrays=[vt], params=[L, vw],
                               use_weights=weighted, seed=3, walker_num=V * 500, max_length=L)
    cnt = tag_transitions(pa, la, vt)
    P = tag_probs(rp, col, w, et, vt, cnt.keys(),
                  lambda v, wt, pt, e: 1 + (pt + wt + e) % TAGS + vt[v] * vw)
    stat, dof, pv = chi2_rows(cnt, P)
    assert dof > 500 and pv > 1e-4, (stat, dof, pv)


@pytest.mark.parametrize("weighted", [True, False])
def test_tag_outlier_walk_follows_its_closed_form(build_dir, weighted):
    """src/tests/test_outlier.cpp:254-390: edges whose tag equals the walker's
    get the dynamic component times the amplifier, above the upper bound; the
    outlier appendix must still give P(e) proportional to static * dynamic."""
    V, E = 40, 140
    rp, col, w, et, vt = tagged_graph(V, E, 22, weighted, sort_by_tag=True)
    begin = np.zeros(V * TAGS, np.uint64)
    end = np.zeros(V * TAGS, np.uint64)
    for v in range(V):
        b, e = int(rp[v]), int(rp[v + 1])
        for t in range(TAGS):
            idx = np.nonzero(et[b:e] == t)[0]
            begin[v * TAGS + t] = b + (idx[0] if len(idx) else 0)
            end[v * TAGS + t] = b + (idx[-1] + 1 if len(idx) else 0)
    amp = 3.0
    L = 20
    pa, la, _ = O.program_walk(plib("tag_outlier", build_dir), rp, col, w, et, arrays=[vt, begin, end],
                               params=[L, amp, 1.0 if weighted else 0.0], use_weights=weighted, seed=4,
                               walker_num=V * 500, max_length=L)
    cnt = tag_transitions(pa, la, vt)
    P = tag_probs(rp, col, w, et, vt, cnt.keys(),
                  lambda v, wt, pt, e: (1 + (pt + wt + e) % TAGS) * (amp if wt == e else 1.0))
    stat, dof, pv = chi2_rows(cnt, P)
    assert dof > 500 and pv > 1e-4, (stat, dof, pv)


def test_return_bias_follows_its_closed_form(build_dir, karate):
    """src/examples/dynamic_walk.cpp: returning to the previous vertex weighs
    twice its static weight."""
    src, dst, w = karate
    rp, col, ww = O.csr_from_edges(34, src, dst, w, False, True)
    L = 20
    pa, la, _ = O.program_walk(plib("return_bias", build_dir), rp, col, ww, params=[L], use_weights=True,
                               seed=6, walker_num=34 * 400, max_length=L)
    cnt = defaultdict(lambda: defaultdict(int))
    for row, n in zip(pa, la):
        r = row[:n].tolist()
        for s in range(1, len(r) - 1):
            cnt[(r[s - 1], r[s])][r[s + 1]] += 1
    P = {}
    for (a, v) in cnt:
        b, e = int(rp[v]), int(rp[v + 1])
        vals = {int(col[k]): float(ww[k]) * (2.0 if int(col[k]) == a else 1.0) for k in range(b, e)}
        tot = sum(vals.values())
        P[(a, v)] = {x: y / tot for x, y in vals.items()}
    stat, dof, pv = chi2_rows(cnt, P)
    assert dof > 100 and pv > 1e-4, (stat, dof, pv)


def test_closed_form_catches_a_planted_mistake(build_dir, tmp_path):
    """The tag walk with its vertex term dropped from the dynamic component
    (bounds unchanged) is rejected by the same check."""
    with open(program_path("tag_walk")) as f:
        src = f.read()
    bad = src.replace("temp += (float)vtag(c)[cur] * c.param[1];", "")
    assert bad != src
    path = tmp_path / "tag_walk_planted.cpp"
    path.write_text(bad)
    V, E = 40, 140
    rp, col, w, et, vt = tagged_graph(V, E, 21, True)
    vw = 0.5
    pa, la, _ = O.program_walk(O.program_oracle(str(path), build_dir), rp, col, w, et, arrays=[vt],
                               params=[20, vw], use_weights=True, seed=3, walker_num=V * 500, max_length=20)
    cnt = tag_transitions(pa, la, vt)
    P = tag_probs(rp, col, w, et, vt, cnt.keys(),
                  lambda v, wt, pt, e: 1 + (pt + wt + e) % TAGS + vt[v] * vw)
    _, _, pv = chi2_rows(cnt, P)
    assert pv < 1e-6


def test_walk_cap_and_second_order_retest(build_dir):
    """max_length cuts a walk that is still going (flag set, length cap + 1);
    a second-order program re-tests the extension on each retry: with
    extension 1 and no dynamic rejection both modes give the same paths."""
    rp, col, w = O.rmat_graph(8, 8, seed=2, weighted=True)
    lib = plib("deepwalk", build_dir)
    pa, la, ta = O.program_walk(lib, rp, col, w, params=[1000], use_weights=True, seed=1, walker_num=500,
                                max_length=12)
    deg = np.diff(rp.astype(np.int64))
    live = deg[np.arange(500) % (len(rp) - 1)] > 0
    assert np.all(la[live] == 13) and np.all(ta[live])
    assert np.all(la[~live] == 1) and not np.any(ta[~live])
    pb, lb, tb = O.program_walk(lib, rp, col, w, params=[1000], use_weights=True, seed=1, walker_num=500,
                                max_length=12, second_order=True)
    same_paths(pa, la, pb, lb)


def test_local_alias_tables_are_the_id_tables():
    """orc_alias_build_local: the same buckets, with local edge indices in
    place of neighbour ids (col[row begin + local] gives the id back)."""
    rp, col, w = O.rmat_graph(11, 16, seed=8, weighted=True)
    p, prim, al, _, vs = O.alias_build(rp, col, w)
    lp, lprim, lal, lvs = O.alias_build_local(rp, w)
    E = len(col)
    begin = np.repeat(rp[:-1].astype(np.int64), np.diff(rp.astype(np.int64)))
    assert np.array_equal(p.view(np.uint32), lp[:E].view(np.uint32))
    assert np.array_equal(prim, col[begin + lprim[:E]])
    assert np.array_equal(al, col[begin + lal[:E]])
    assert np.array_equal(vs.view(np.uint32), lvs.view(np.uint32))
