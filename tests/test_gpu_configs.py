"""The benchmark's full-size workloads against the oracle.

node2vec p=2 q=0.5 on the undirected R-MAT scale-24 graph (16,777,216
vertices, ~521 M directed edges), weighted and unweighted, one walker per
vertex, L = 80, in both launch forms: full-size properties on every walker,
the alias tables bit for bit against the oracle's own construction, and the
rows of every 64th walker bit for bit against the oracle's walk.  PPR serving
on the same resident graph: a Poisson stream of single-source queries, every
served result against the oracle's walk of that query, and the pooled path
lengths against the geometric law.  (Reference apps: src/apps/node2vec.hpp:108-197,
src/apps/ppr.hpp:30-64.)
"""
import time

import numpy as np
import pytest

from helpers import (O, assert_rows_match_oracle, edge_keys, edges_exist, kk, oracle_walk,
                     packed_matches_oracle, rows_equal)

pytestmark = pytest.mark.gpu

SCALE = 24
SAMPLE_STRIDE = 64


@pytest.fixture(scope="module")
def rmat24(engine):
    print(f"[gpu] R-MAT scale {SCALE}, weighted, seed 1 (the benchmark graph)", flush=True)
    engine.graph_rmat(SCALE, 16, 0.57, 0.19, 0.19, seed=1, weighted=True)
    engine.prepare()
    rp, col, w, _ = engine.download()
    g = O.Graph(rp, col, w)
    yield g
    g.alias = None


def test_rmat24_graph_and_alias_tables(engine, rmat24):
    """CSR shape (sorted lists, symmetric degrees) and the alias tables against
    the oracle's sequential FIFO construction, every bucket bit for bit."""
    g = rmat24
    rp, col = g.row_ptr, g.col
    assert g.V == 1 << SCALE and rp[0] == 0 and int(rp[-1]) == len(col)
    deg = np.diff(rp.astype(np.int64))
    # lists sorted and free of duplicates and self loops
    d = np.diff(col.astype(np.int64))
    first = np.zeros(len(col), bool)
    first[rp[:-1][deg > 0].astype(np.int64)] = True
    assert np.all((d > 0) | first[1:])
    src = np.repeat(np.arange(g.V, dtype=np.uint32), deg)
    assert not np.any(src == col)
    # mirrored: the in-degree of every vertex equals its out-degree
    assert np.array_equal(np.bincount(col, minlength=g.V), deg)
    del src, d, first
    print("[gpu] alias tables: device download vs oracle construction", flush=True)
    p, prim, al, meta, vs = engine.alias_download()
    g.with_alias()
    p_o, prim_o, al_o, meta_o, vs_o = g.alias
    assert np.array_equal(vs.view(np.uint32), vs_o.view(np.uint32))
    assert np.array_equal(p.view(np.uint32), p_o.view(np.uint32))
    assert np.array_equal(prim, prim_o)
    assert np.array_equal(al, al_o)


@pytest.mark.parametrize("weighted", [True, False], ids=["weighted", "unweighted"])
def test_node2vec_rmat24_config(engine, rmat24, weighted):
    g = rmat24
    V = g.V
    deg = np.diff(g.row_ptr.astype(np.int64))
    job = kk.node2vec_job(V, 80, 2.0, 0.5, weighted, seed=1)
    print(f"[gpu] node2vec R-MAT-24 weighted={weighted}: persistent launch, {V} walkers", flush=True)
    (pp,) = engine.run([job], mode=kk.MODE_PERSISTENT)
    # every walker from a non-isolated vertex takes 80 steps; isolated ones none
    assert np.array_equal(pp.lengths, np.where(deg > 0, 81, 1))
    assert not pp.truncated.any()
    assert np.array_equal(pp.vertices[:, 0], np.arange(V, dtype=np.uint32))
    # bit for bit against the oracle on every 64th walker
    t0 = time.perf_counter()
    n = V // SAMPLE_STRIDE
    op, ol, ot, st = O.walk(g, O.NODE2VEC, n, 80, weighted, p=2.0, q=0.5, seed=1, walker_stride=SAMPLE_STRIDE)
    print(f"[gpu] oracle sample: {n} walkers, {st.steps} steps in {time.perf_counter() - t0:.1f}s", flush=True)
    rows = np.arange(n, dtype=np.int64) * SAMPLE_STRIDE
    assert_rows_match_oracle(pp, op, ol, ot, rows, f"node2vec R-MAT-24 weighted={weighted} (sample)")
    # transitions of other walkers follow CSR edges
    rng = np.random.default_rng(3)
    other = rng.choice(np.nonzero(deg > 0)[0], 20000, replace=False)
    assert edges_exist(edge_keys(g.row_ptr, g.col), pp.vertices[other, :80].reshape(-1),
                       pp.vertices[other, 1:81].reshape(-1))
    # the plain launch gives every row of the persistent one
    print(f"[gpu] node2vec R-MAT-24 weighted={weighted}: plain launch", flush=True)
    (pl,) = engine.run([job], mode=kk.MODE_PLAIN)
    assert rows_equal(pl.vertices, pl.lengths, pp.vertices, pp.lengths), "plain and persistent rows differ"


def _geometric_chi2(steps, t, k_max=40):
    """Pooled chi-square of path step counts against P(k) = (1-t)^k t (tail
    from k_max pooled)."""
    from scipy import stats
    ext = float(np.float32(1.0) - np.float32(t))
    # continue iff u01 <= ext with u01 = j * 2^-24: P(continue) = (floor(ext 2^24) + 1) / 2^24
    pc = (np.floor(ext * 2.0 ** 24) + 1) / 2.0 ** 24
    n = len(steps)
    obs = np.bincount(np.minimum(steps, k_max), minlength=k_max + 1).astype(np.float64)
    k = np.arange(k_max)
    exp = np.concatenate([n * pc ** k * (1 - pc), [n * pc ** k_max]])
    stat = float(np.sum((obs - exp) ** 2 / exp))
    return stat, k_max, stats.chi2.sf(stat, k_max)


def test_ppr_serving_rmat24(engine, rmat24):
    """Poisson PPR queries (t = 0.15, 10^3..10^5 walkers log-uniform, source =
    the source of a uniformly drawn directed edge) at 1000 queries/s on the
    resident R-MAT-24 graph: every result equals the oracle's walk of that
    query, and the pooled lengths follow the geometric law."""
    g = rmat24
    rp = g.row_ptr
    E = int(rp[-1])
    rng = np.random.default_rng(17)
    n_q, rate, t = 400, 1000.0, 0.15
    gaps = rng.exponential(1.0 / rate, n_q)
    srcs = np.searchsorted(rp, rng.integers(0, E, n_q), side="right") - 1
    walkers = np.round(10 ** (3 + 2 * rng.random(n_q))).astype(np.int64)
    jobs = [kk.ppr_job(int(walkers[i]), t, True, seed=5000 + i, start_vertices=np.array([srcs[i]], np.uint32))
            for i in range(n_q)]
    print(f"[gpu] serving {n_q} PPR queries at {rate:.0f}/s on R-MAT-24, {walkers.sum()} walkers", flush=True)
    engine.serve_start(max_walkers_in_flight=1 << 22)
    try:
        tickets = []
        t_next = time.perf_counter()
        for i in range(n_q):
            t_next += gaps[i]
            while time.perf_counter() < t_next:
                pass
            tickets.append(engine.submit(jobs[i]))
        results = [engine.wait(tk, 60.0, packed=True) for tk in tickets]
    finally:
        engine.serve_stop()
    lat = np.array([ms for _, ms in results])
    print(f"[gpu] latency p50 {np.percentile(lat, 50):.3f} ms p99 {np.percentile(lat, 99):.3f} ms", flush=True)
    all_steps = []
    for i, ((got, _), job) in enumerate(zip(results, jobs)):
        op, ol, ot, _ = oracle_walk(g, job)
        assert packed_matches_oracle(got, op, ol, ot), f"query {i} (source {srcs[i]}, seed {job.seed}) differs"
        all_steps.append(ol.astype(np.int64) - 1)
    steps = np.concatenate(all_steps)
    stat, dof, pv = _geometric_chi2(steps, t)
    print(f"[gpu] geometric law over {len(steps)} walkers: chi2 {stat:.1f} on {dof} dof, p={pv:.3g}", flush=True)
    assert pv > 1e-4
    # a planted mistake in the law (t taken as 0.16) is rejected
    assert _geometric_chi2(steps, 0.16)[2] < 1e-6
