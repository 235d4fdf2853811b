"""Walk programs on the GPU (kkw_program_*): compiled at run time with hipRTC,
run through the C ABI, bit-exact against the oracle's program loop compiled
with the same source; programs restating the built-in apps are also
bit-exact against the built-in kernels."""
import os

import numpy as np
import pytest

import knightking_amd as kk
from oracle import oracle as O
from test_oracle_programs import TAGS, program_path, same_paths, tagged_graph

pytestmark = pytest.mark.gpu

# both launch forms of a program: one lane per walker, and lanes refilled from a work queue
MODES = [pytest.param(kk.MODE_PLAIN, id="plain"), pytest.param(kk.MODE_PERSISTENT, id="persistent")]
# tables of the app programs: the engine's (ids + rows; the programs' own
# KKW_EDGE_INDEX 0) and local-index tables (rebuilt with KKW_EDGE_INDEX 1)
TABLES = [pytest.param(None, id="id_tables"), pytest.param("-DKKW_EDGE_INDEX=1", id="index_tables")]


@pytest.fixture(scope="module")
def build_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("program_oracles"))


def source(name):
    with open(program_path(name)) as f:
        return f.read()


def both(engine, build_dir, name, rp, col, w=None, et=None, arrays=(), options=None, **run):
    """Run program `name` on the GPU and on the oracle; return both results."""
    print(f"[gpu] program {name} V={len(rp) - 1} E={len(col)} options={options} {run}", flush=True)
    engine.graph_from_csr(rp, col, w)
    if et is not None:
        engine.set_edge_types(et)
    prog = engine.program(source(name), options)
    try:
        for k, a in enumerate(arrays):
            prog.set_array(k, a)
        got = prog.run(**run)
    finally:
        prog.close()
    want = O.program_walk(O.program_oracle(program_path(name), build_dir), rp, col, w, et,
                          arrays=arrays, params=run.get("params", ()), use_weights=run.get("use_weights", False),
                          seed=run.get("seed", 1), walker_num=run["walker_num"], max_length=run["max_length"],
                          walker_begin=run.get("walker_begin", 0), start_vertices=run.get("start_vertices"),
                          uniform_start=run.get("uniform_start", False),
                          second_order=run.get("second_order", False))
    return got, want


def check(got, want):
    pw, lw, tw = want
    same_paths(got.vertices, got.lengths, pw, lw)
    assert np.array_equal(got.truncated, tw)
    assert got.steps == int((lw.astype(np.int64) - 1).sum())


@pytest.mark.parametrize("tables", TABLES)
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("weighted,L", [(True, 40), (False, 40), (True, 80)])
def test_deepwalk_program(engine, build_dir, weighted, L, mode, tables):
    # L = 80: persistent rows of whole lines (half-line write-out from LDS)
    rp, col, w = O.rmat_graph(12, 16, seed=3, weighted=True)
    check(*both(engine, build_dir, "deepwalk", rp, col, w, walker_num=20000, max_length=L, params=[L],
                use_weights=weighted, seed=9, mode=mode, options=tables))


@pytest.mark.parametrize("tables", TABLES)
@pytest.mark.parametrize("mode", MODES)
def test_ppr_program_uniform_start(engine, build_dir, mode, tables):
    rp, col, w = O.rmat_graph(12, 16, seed=4, weighted=True)
    t = 0.15
    ext = np.float32(1.0) - np.float32(t)
    check(*both(engine, build_dir, "ppr", rp, col, w, walker_num=30000, max_length=O.ppr_default_cap(t),
                params=[ext], use_weights=True, seed=2, uniform_start=True, walker_begin=77, mode=mode,
                options=tables))


@pytest.mark.parametrize("tables", TABLES)
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("p,q,weighted", [(2.0, 0.5, True), (0.5, 2.0, True), (0.25, 1.0, False)])
def test_node2vec_program_matches_oracle_and_builtin_kernel(engine, build_dir, p, q, weighted, mode, tables):
    rp, col, w = O.rmat_graph(12, 16, seed=5, weighted=True)
    L = 40
    up, lo, ip, iq, ov = kk.node2vec_constants(p, q)
    got, want = both(engine, build_dir, "node2vec", rp, col, w, walker_num=8192, max_length=L,
                     params=[L, up, lo, ip, iq, ov, 1.0 if weighted else 0.0], use_weights=weighted, seed=11,
                     second_order=True, mode=mode, options=tables)
    check(got, want)
    (builtin,) = engine.run([kk.node2vec_job(8192, L, p, q, weighted, seed=11)], mode=kk.MODE_PLAIN)
    same_paths(got.vertices, got.lengths, builtin.vertices, builtin.lengths)


@pytest.mark.parametrize("tables", TABLES)
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("p,q,weighted", [(2.0, 0.5, True), (0.5, 2.0, True), (0.25, 1.0, False)])
def test_node2vec_query_program_matches_oracle_and_builtin_kernel(engine, build_dir, p, q, weighted, mode, tables):
    """node2vec through the second-order query / response functions: the
    device loop against the oracle's loop and the built-in kernel."""
    rp, col, w = O.rmat_graph(12, 16, seed=5, weighted=True)
    L = 40
    up, lo, ip, iq, ov = kk.node2vec_constants(p, q)
    got, want = both(engine, build_dir, "node2vec_query", rp, col, w, walker_num=8192, max_length=L,
                     params=[L, up, lo, ip, iq, ov, 1.0 if weighted else 0.0], use_weights=weighted, seed=11,
                     mode=mode, options=tables)
    check(got, want)
    (builtin,) = engine.run([kk.node2vec_job(8192, L, p, q, weighted, seed=11)], mode=kk.MODE_PLAIN)
    same_paths(got.vertices, got.lengths, builtin.vertices, builtin.lengths)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("weighted", [True, False])
def test_second_order_answers_kept_between_queries(engine, build_dir, weighted, mode):
    """A response carrying data, queries sent on even steps only, the kept
    answer seen on odd ones: device loop and oracle loop agree bit for bit."""
    rp, col, w = O.rmat_graph(11, 16, seed=8, weighted=True)
    check(*both(engine, build_dir, "edge_weight_query", rp, col, w, walker_num=10000, max_length=30, params=[30],
                use_weights=weighted, seed=21, mode=mode))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("weighted", [True, False])
def test_tag_walk_program(engine, build_dir, weighted, mode):
    rp, col, w, et, vt = tagged_graph(300, 2000, 21, weighted)
    check(*both(engine, build_dir, "tag_walk", rp, col, w, et, arrays=[vt], walker_num=20000, max_length=20,
                params=[20, 0.5], use_weights=weighted, seed=3, mode=mode))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("weighted", [True, False])
def test_tag_outlier_program(engine, build_dir, weighted, mode):
    V = 300
    rp, col, w, et, vt = tagged_graph(V, 2000, 22, weighted, sort_by_tag=True)
    begin = np.zeros(V * TAGS, np.uint64)
    end = np.zeros(V * TAGS, np.uint64)
    for v in range(V):
        b, e = int(rp[v]), int(rp[v + 1])
        for t in range(TAGS):
            idx = np.nonzero(et[b:e] == t)[0]
            begin[v * TAGS + t] = b + (idx[0] if len(idx) else 0)
            end[v * TAGS + t] = b + (idx[-1] + 1 if len(idx) else 0)
    check(*both(engine, build_dir, "tag_outlier", rp, col, w, et, arrays=[vt, begin, end], walker_num=20000,
                max_length=20, params=[20, 3.0, 1.0 if weighted else 0.0], use_weights=weighted, seed=4, mode=mode))


@pytest.mark.parametrize("mode", MODES)
def test_return_bias_program_start_list_and_cap(engine, build_dir, karate, mode):
    src, dst, w = karate
    rp, col, ww = O.csr_from_edges(34, src, dst, w, False, True)
    check(*both(engine, build_dir, "return_bias", rp, col, ww, walker_num=5000, max_length=7,
                params=[1000], use_weights=True, seed=6, start_vertices=np.array([0, 33, 5], np.uint32), mode=mode))


# A static draw of a program built with KKW_EDGE_INDEX 0 carries
# KKW_NO_EDGE_INDEX; an outlier edge keeps its index.  The walker stops at
# the first step whose edge broke that rule, so full-length paths say it held.
EDGE_INDEX_PROBE = """
#define KKW_EDGE_INDEX 0
#define KKW_DYNAMIC 1
#define KKW_OUTLIER 1
struct WalkerState { uint32_t ok; };
KKW_FN void init_state(WalkerState &s, const kkw_walker &, uint32_t, const kkw_ctx &) { s.ok = 1; }
KKW_FN void update_state(WalkerState &s, const kkw_walker &, uint32_t cur, const kkw_edge &e, const kkw_ctx &c)
{
    const bool outlier = e.index != KKW_NO_EDGE_INDEX;
    if (outlier && (e.index != c.row_ptr[cur] || e.neighbour != c.col[e.index])) s.ok = 0;
}
KKW_FN float extension(const WalkerState &s, const kkw_walker &w, uint32_t, const kkw_ctx &c)
{
    return (s.ok && w.step < (uint32_t)c.param[0]) ? 1.0f : 0.0f;
}
KKW_FN float dynamic_comp(const WalkerState &, const kkw_walker &, uint32_t, const kkw_edge &e, const kkw_ctx &)
{
    return e.index != KKW_NO_EDGE_INDEX ? 1e30f : 1.0f;   // outlier darts reach above the upper bound
}
KKW_FN float upper_bound(uint32_t, const kkw_ctx &) { return 1.0f; }
KKW_FN void outlier_bound(const WalkerState &, const kkw_walker &, uint32_t, float &overflow, uint32_t &num,
                          const kkw_ctx &)
{
    overflow = 0.5f;
    num = 1;
}
KKW_FN int64_t outlier_search(const WalkerState &, const kkw_walker &, uint32_t cur, uint32_t, const kkw_ctx &c)
{
    return (int64_t)c.row_ptr[cur];   // the first edge of the list
}
"""


@pytest.mark.parametrize("weighted", [True, False])
def test_edge_index_sentinel_of_id_tables(engine, weighted):
    rp, col, w = O.rmat_graph(10, 8, seed=7, weighted=True)
    engine.graph_from_csr(rp, col, w)
    prog = engine.program(EDGE_INDEX_PROBE)
    try:
        got = prog.run(4096, 20, params=[20], use_weights=weighted, seed=3, mode=kk.MODE_PERSISTENT)
    finally:
        prog.close()
    deg = np.diff(rp.astype(np.int64))
    starts = np.arange(4096) % (len(rp) - 1)
    live = deg[starts] > 0
    assert np.all(got.lengths[live] == 21)   # every step kept the rule
    assert np.all(got.lengths[~live] == 1)


def test_program_errors(engine):
    rp, col, w = O.rmat_graph(8, 8, seed=2, weighted=True)
    engine.graph_from_csr(rp, col, w)
    with pytest.raises(kk.KkwError) as e:
        engine.program("struct WalkerState { int x; };\nthis does not compile\n")
    assert e.value.code == kk.KKW_E_INVALID and "compile" in str(e.value)
    prog = engine.program(source("deepwalk"))
    try:
        with pytest.raises(kk.KkwError) as e:
            prog.set_array(8, np.zeros(4, np.int32))
        assert e.value.code == kk.KKW_E_INVALID
        with pytest.raises(kk.KkwError) as e:
            prog.run(10, 5, params=[5], start_vertices=np.array([1 << 20], np.uint32))
        assert e.value.code == kk.KKW_E_INVALID
        got = prog.run(0, 5, params=[5])
        assert got.lengths.shape == (0,)
        got = prog.run(0, 5, params=[5], mode=kk.MODE_PERSISTENT)
        assert got.lengths.shape == (0,)
        with pytest.raises(kk.KkwError) as e:
            prog.run(10, 5, params=[5], mode=7)
        assert e.value.code == kk.KKW_E_INVALID
    finally:
        prog.close()


@pytest.mark.parametrize("mode", MODES)
def test_envelope_violations_are_counted(engine, mode):
    """kkw_timing.envelope_violations (the reference's check at
    include/walk.hpp:1003-1005): 0 for correct programs -- the return-biased
    walk (upper bound 2) and node2vec with its outlier appendix (p = 0.5) --
    and the number of offending trials when the upper bound is set too low
    (1.5 under a dynamic component of 2 for the return edge)."""
    rp, col, w = O.rmat_graph(10, 16, seed=6, weighted=True)
    engine.graph_from_csr(rp, col, w)
    src = source("return_bias")
    prog = engine.program(src)
    try:
        prog.run(4000, 20, use_weights=True, seed=3, params=[20], mode=mode)
        assert engine.timing()["envelope_violations"] == 0
    finally:
        prog.close()
    up, lo, ip, iq, ov = kk.node2vec_constants(0.5, 2.0)
    prog = engine.program(source("node2vec"))
    try:
        prog.run(4000, 20, use_weights=True, seed=3, params=[20, up, lo, ip, iq, ov, 1.0], second_order=True,
                 mode=mode)
        assert engine.timing()["envelope_violations"] == 0
    finally:
        prog.close()
    bad = src.replace("KKW_FN float upper_bound(uint32_t, const kkw_ctx &) { return 2.0f; }",
                      "KKW_FN float upper_bound(uint32_t, const kkw_ctx &) { return 1.5f; }")
    assert bad != src
    prog = engine.program(bad)
    try:
        got = prog.run(4000, 20, use_weights=True, seed=3, params=[20], mode=mode)
        n = engine.timing()["envelope_violations"]
        # every accepted or rejected trial of the return edge after step 0 is one
        rows = got.vertices[:, :21]
        returns = 0
        for i in range(4000):
            n_i = int(got.lengths[i])
            if n_i >= 3:
                returns += int(np.sum(rows[i, 2:n_i] == rows[i, :n_i - 2]))
        assert n >= returns > 0
    finally:
        prog.close()
