"""Serving loop: continuous batching of walk queries on a resident graph.

A query's paths are the same whether it runs alone or admitted into a mix
with others (draws are keyed by seed and walker id, not by slot or time).
Failure handling is forced from the host: an admission budget smaller than a
request, a full queue, a cancelled request, malformed descriptors.
"""
import threading

import numpy as np
import pytest

import knightking_amd as kk

pytestmark = pytest.mark.gpu


def same(a, b):
    """a: Paths or PackedPaths; b: Paths (rows of `stride` entries)."""
    if not np.array_equal(a.lengths, b.lengths):
        return False
    if isinstance(a, kk.PackedPaths):
        mask = np.arange(b.vertices.shape[1])[None, :] < b.lengths[:, None]
        return bool(np.array_equal(a.vertices, b.vertices[mask]) and a.offsets[-1] == len(a.vertices))
    # row strides may differ (persistent batches pad rows to whole 128-byte lines)
    n = min(a.vertices.shape[1], b.vertices.shape[1])
    mask = np.arange(n)[None, :] < a.lengths[:, None]
    return bool(np.all((a.vertices[:, :n] == b.vertices[:, :n]) | ~mask))


@pytest.fixture()
def served(engine):
    print("[gpu] serving on R-MAT scale 12", flush=True)
    engine.graph_rmat(12, 16, seed=3, weighted=True)
    engine.type_by_dst(5, 4)
    engine.prepare()
    yield engine
    engine.serve_stop()


def mixed_queries(V, n, seed):
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        k = i % 4
        walkers = int(10 ** (2 + 2 * rng.random()))
        if k == 0:
            jobs.append(kk.ppr_job(walkers, 0.15, True, seed=100 + i,
                                   start_vertices=np.array([rng.integers(0, V)], np.uint32)))
        elif k == 1:
            jobs.append(kk.deepwalk_job(walkers, int(rng.integers(5, 40)), bool(i % 2), seed=100 + i))
        elif k == 2:
            jobs.append(kk.node2vec_job(walkers, 20, 2.0, 0.5, True, seed=100 + i))
        else:
            jobs.append(kk.metapath_job(walkers, 15, [[0b00111, 0b11000], [0b11111]], True, seed=100 + i))
    return jobs


def test_results_same_alone_and_in_the_mix(served):
    """Queries of every kind submitted from four threads: each served result
    (packed or row form, through flights, packing and the pinned hand-back)
    equals the oracle's walk of that query and the query run alone."""
    from helpers import assert_rows_match_oracle, oracle_walk, packed_matches_oracle
    from oracle import oracle as O
    eng = served
    V, _ = eng.info()
    rp, col, w, et = eng.download(etype=True)
    g = O.Graph(rp, col, w, et)
    jobs = mixed_queries(V, 40, 1)
    alone = [eng.run([j])[0] for j in jobs]
    eng.serve_start(max_walkers_in_flight=20000)
    tickets = [None] * len(jobs)

    def submit(idx):
        for i in idx:
            tickets[i] = eng.submit(jobs[i])

    threads = [threading.Thread(target=submit, args=(range(t, len(jobs), 4),)) for t in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for i, t in enumerate(tickets):
        got, lat = eng.wait(t, 60.0, packed=bool(i % 2))
        assert lat >= 0
        op, ol, ot, _ = oracle_walk(g, jobs[i])
        if i % 2:
            assert packed_matches_oracle(got, op, ol, ot), f"query {i}: served result differs from the oracle"
        else:
            assert_rows_match_oracle(got, op, ol, ot, what=f"query {i} served")
        assert same(got, alone[i]), f"query {i} differs in the mix"


def test_idle_polling_serves_the_same_paths(served):
    """With idle_poll_us the idle loop keeps looking for queries instead of
    sleeping: queries spaced so that the loop goes idle before each (some
    inside its polling window, one after it) come back equal to the oracle's
    walks, and the loop stops while it polls."""
    import time
    from helpers import assert_rows_match_oracle, oracle_walk
    from oracle import oracle as O
    eng = served
    V, _ = eng.info()
    rp, col, w, et = eng.download(etype=True)
    g = O.Graph(rp, col, w, et)
    jobs = mixed_queries(V, 8, 5)
    eng.serve_start(max_walkers_in_flight=20000, idle_poll_us=3000)
    for i, job in enumerate(jobs):
        t = eng.submit(job)
        got, lat = eng.wait(t, 60.0)
        assert lat >= 0
        op, ol, ot, _ = oracle_walk(g, job)
        assert_rows_match_oracle(got, op, ol, ot, what=f"query {i} served with idle polling")
        time.sleep(0.005 if i == 3 else 0.001)   # once past the 3 ms window, else inside it
    eng.serve_stop()


def test_budget_smaller_than_request_fails_it(engine):
    eng = kk.Engine(hbm_budget=1 << 20)
    try:
        eng.graph_rmat(10, 16, seed=1, weighted=False)
        eng.serve_start()
        big = eng.submit(kk.deepwalk_job(100000, 80, False, seed=1))   # ~32 MB of paths
        small = eng.submit(kk.deepwalk_job(100, 10, False, seed=2))
        with pytest.raises(kk.KkwError) as e:
            eng.wait(big, 30.0)
        assert e.value.code == kk.KKW_E_NOMEM
        got, _ = eng.wait(small, 30.0)
        assert got.lengths.shape == (100,)
    finally:
        eng.serve_stop()
        eng.close()


def test_queue_full_gives_back_pressure(served):
    eng = served
    eng.serve_start(max_walkers_in_flight=1 << 20, queue_capacity=3)
    accepted, busy = [], 0
    for i in range(60):
        try:
            accepted.append(eng.submit(kk.deepwalk_job(1 << 20, 80, True, seed=i)))
        except kk.KkwError as e:
            assert e.code == kk.KKW_E_BUSY
            busy += 1
    assert busy > 0
    for t in accepted:
        got, _ = eng.wait(t, 120.0)
        assert got.lengths.shape == (1 << 20,)


def test_cancel_pending_request(served):
    eng = served
    eng.serve_start(max_walkers_in_flight=1 << 21, queue_capacity=64)
    first = eng.submit(kk.deepwalk_job(1 << 21, 80, True, seed=1))
    second = eng.submit(kk.deepwalk_job(1 << 21, 80, True, seed=2))
    third = eng.submit(kk.deepwalk_job(1000, 10, True, seed=3))
    eng.cancel(third)
    codes = []
    for t in (first, second, third):
        try:
            eng.wait(t, 120.0)
            codes.append(0)
        except kk.KkwError as e:
            codes.append(e.code)
    assert codes[0] == 0 and codes[1] == 0
    assert codes[2] in (0, kk.KKW_E_CANCELLED)


@pytest.mark.parametrize("bad", ["start", "scheme", "kind", "pq", "ppr_t"])
def test_malformed_descriptor_rejected_at_submit(served, bad):
    eng = served
    V, _ = eng.info()
    eng.serve_start()
    if bad == "start":
        job = kk.ppr_job(10, 0.15, True, start_vertices=np.array([V + 5], np.uint32))
    elif bad == "scheme":
        job = kk.metapath_job(10, 10, [], True)
    elif bad == "kind":
        job = kk.Job(7, 10, 10)
    elif bad == "pq":
        job = kk.node2vec_job(10, 10, 0.0, 1.0)
    else:
        job = kk.ppr_job(10, 1.5, True)
    with pytest.raises(kk.KkwError) as e:
        eng.submit(job)
    assert e.value.code == kk.KKW_E_INVALID


def test_results_collected_out_of_order_under_pinned_pressure(engine):
    """Served results live in pinned memory until collected; with a pinned
    arena that holds only a few of them and results collected newest first,
    admission must still progress (finished results are moved to heap memory)
    and every result must equal the same job run alone."""
    eng = kk.Engine(hbm_budget=96 << 20)   # device and pinned arenas of 96 MiB
    try:
        eng.graph_rmat(12, 16, seed=3, weighted=True)
        eng.prepare()
        jobs = [kk.deepwalk_job(20000, 80, True, seed=500 + i) for i in range(40)]   # ~6.7 MB of paths each
        print("[gpu] 40 served DeepWalk queries, collected newest first", flush=True)
        eng.serve_start(max_walkers_in_flight=1 << 20)
        tickets = [eng.submit(j) for j in jobs]
        got = {}
        for i in reversed(range(len(jobs))):
            got[i], _ = eng.wait(tickets[i], 60.0)
        eng.serve_stop()
        for i in (0, 13, 39):
            (alone,) = eng.run([jobs[i]])
            assert same(got[i], alone), f"query {i} differs"
    finally:
        eng.serve_stop()
        eng.close()


def test_graph_cannot_change_while_serving(served):
    """Entry points that replace or free device tables refuse with KKW_E_BUSY
    while the loop serves (its flights read the tables without the engine
    lock); queries keep their answers, and the calls work again after stop."""
    eng = served
    V, E = eng.info()
    job = kk.node2vec_job(3000, 20, 2.0, 0.5, True, seed=9)
    (ref,) = eng.run([job])
    eng.serve_start(max_walkers_in_flight=1 << 16)
    t = eng.submit(job)
    changes = {
        "graph_rmat": lambda: eng.graph_rmat(8, 8, seed=1, weighted=True),
        "graph_from_edges": lambda: eng.graph_from_edges(4, np.array([0, 1], np.uint32), np.array([1, 2], np.uint32)),
        "prepare": eng.prepare,
        "adj_index_off": lambda: eng.adj_index(False),
        "type_by_dst": lambda: eng.type_by_dst(3, 1),
        "set_edge_types": lambda: eng.set_edge_types(np.zeros(E, np.uint8)),
        "sort_adj": eng.sort_adj,
    }
    for name, call in changes.items():
        with pytest.raises(kk.KkwError) as e:
            call()
        assert e.value.code == kk.KKW_E_BUSY, name
    got, _ = eng.wait(t, 60.0)
    assert same(got, ref)
    eng.serve_stop()
    eng.adj_index(False)   # allowed again
    eng.adj_index(True)
    assert eng.info() == (V, E)


@pytest.mark.parametrize("round_steps", [1, 3, 16])
def test_rounds_give_the_same_paths(served, round_steps):
    """Flights walked in rounds of at most round_steps steps per walker
    (walkers suspended at step boundaries and continued by the next launch):
    every kind, every result equal to the oracle's walk of that query."""
    from helpers import oracle_walk, packed_matches_oracle
    from oracle import oracle as O
    eng = served
    V, _ = eng.info()
    rp, col, w, et = eng.download(etype=True)
    g = O.Graph(rp, col, w, et)
    jobs = mixed_queries(V, 24, 5)
    eng.serve_start(max_walkers_in_flight=20000, round_steps=round_steps)
    tickets = [eng.submit(j) for j in jobs]
    for i, t in enumerate(tickets):
        got, _ = eng.wait(t, 60.0, packed=True)
        op, ol, ot, _ = oracle_walk(g, jobs[i])
        assert packed_matches_oracle(got, op, ol, ot), f"query {i} differs with rounds of {round_steps}"


def test_rounds_let_a_short_query_pass_a_long_one(engine):
    """A long node2vec query (2^20 walkers, L = 80, R-MAT scale 20) served in
    rounds of 4 steps gives the GPU back between rounds: a short PPR query
    submitted after it completes while the long one is still walking."""
    print("[gpu] rounds: long node2vec query and a short PPR query on R-MAT-20", flush=True)
    engine.graph_rmat(20, 16, seed=2, weighted=True)
    engine.prepare()
    import time
    long_job = kk.node2vec_job(1 << 20, 80, 2.0, 0.5, True, seed=3)
    short_job = kk.ppr_job(1000, 0.15, True, seed=4, start_vertices=np.array([1], np.uint32))
    engine.serve_start(round_steps=4)
    try:
        t_long = engine.submit(long_job)
        time.sleep(0.001)
        t_short = engine.submit(short_job)
        got, lat_short = engine.wait(t_short, 60.0, packed=True)
        long_still_running = not engine.poll(t_long)
        got_long, lat_long = engine.wait(t_long, 60.0, packed=True)
        print(f"[gpu] short {lat_short:.3f} ms, long {lat_long:.3f} ms", flush=True)
        assert got.lengths.shape == (1000,) and got_long.lengths.shape == (1 << 20,)
        assert long_still_running, (lat_short, lat_long)
    finally:
        engine.serve_stop()
