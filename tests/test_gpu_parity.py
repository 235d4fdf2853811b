"""GPU parity: the HIP path through the C ABI against the oracle, bit for bit.

Graph construction (R-MAT, edge lists), alias tables and every walk kind in
both launch forms are compared with the CPU restatement on the same seeded
inputs.  Larger sizes are covered by size-independent properties.
"""
import os

import numpy as np
import pytest

import knightking_amd as kk
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# A persistent launch whose node2vec or DeepWalk job runs in the state-machine
# kernel whatever its size (the library picks that kernel from 16384 walkers
# per CU up, the trial-loop kernel below; other kinds run the trial loop).
PERSISTENT_SM = "persistent_sm"
MODES = [kk.MODE_PLAIN, kk.MODE_PERSISTENT, PERSISTENT_SM]
MODE_IDS = ["plain", "persistent", "persistent_sm"]


def run_jobs(engine, jobs, mode, **kw):
    """engine.run in one of MODES."""
    if mode != PERSISTENT_SM:
        return engine.run(jobs, mode=mode, **kw)
    old = os.environ.get("KKW_WALK_KERNEL")
    os.environ["KKW_WALK_KERNEL"] = "sm"
    try:
        return engine.run(jobs, mode=kk.MODE_PERSISTENT, **kw)
    finally:
        if old is None:
            del os.environ["KKW_WALK_KERNEL"]
        else:
            os.environ["KKW_WALK_KERNEL"] = old


def assert_paths_equal(gpu, o_paths, o_lens, o_trunc, what=""):
    assert np.array_equal(gpu.lengths, o_lens), f"{what}: lengths differ"
    assert np.array_equal(gpu.truncated, o_trunc), f"{what}: truncation flags differ"
    n, stride = o_paths.shape
    # device rows hold whole groups of four entries: at least L+1
    assert gpu.vertices.shape[0] == n and gpu.vertices.shape[1] >= stride
    assert gpu.vertices.shape[1] % 4 == 0
    mask = np.arange(stride)[None, :] < o_lens[:, None]
    bad = np.nonzero((gpu.vertices[:, :stride] != o_paths) & mask)
    assert bad[0].size == 0, f"{what}: {bad[0].size} path entries differ, first at walker {bad[0][:1]}"


def assert_same_rows(a, b, what=""):
    """Two GPU results hold the same paths.  Row strides may differ (persistent
    batches pad rows to whole 128-byte lines); entries past a length do not count."""
    assert np.array_equal(a.lengths, b.lengths), f"{what}: lengths differ"
    n = min(a.vertices.shape[1], b.vertices.shape[1])
    assert a.lengths.size == 0 or int(a.lengths.max()) <= n
    mask = np.arange(n)[None, :] < a.lengths[:, None]
    assert np.array_equal(np.where(mask, a.vertices[:, :n], 0), np.where(mask, b.vertices[:, :n], 0)), what


@pytest.fixture(scope="module")
def rmat10(engine):
    g = O.rmat_graph(10, 16, seed=1, weighted=True)
    return O.Graph(*g)


def load_rmat(engine, scale, seed=1, weighted=True):
    print(f"[gpu] R-MAT scale {scale} seed {seed} weighted={weighted}", flush=True)
    engine.graph_rmat(scale, 16, seed=seed, weighted=weighted)


@pytest.mark.parametrize("scale", [6, 10, 14])
def test_rmat_csr_matches_oracle(engine, scale):
    load_rmat(engine, scale)
    rp, col, w, _ = engine.download()
    rp_o, col_o, w_o = O.rmat_graph(scale, 16, seed=1, weighted=True)
    assert np.array_equal(rp, rp_o)
    assert np.array_equal(col, col_o)
    assert np.array_equal(w.view(np.uint32), w_o.view(np.uint32))


@pytest.mark.parametrize("undirected,sort_adj", [(False, False), (False, True), (True, False), (True, True)])
def test_edges_csr_matches_oracle(engine, karate, undirected, sort_adj):
    src, dst, w = karate
    if undirected:  # keep one direction of each pair, let the loader mirror it
        keep = src < dst
        src, dst, w = src[keep], dst[keep], w[keep]
    print(f"[gpu] karate CSR undirected={undirected} sort_adj={sort_adj}", flush=True)
    engine.graph_from_edges(34, src, dst, w, make_undirected=undirected, sort_adj=sort_adj)
    rp, col, wd, _ = engine.download()
    rp_o, col_o, w_o = O.csr_from_edges(34, src, dst, w, undirected=undirected, sort_adj=sort_adj)
    assert np.array_equal(rp, rp_o) and np.array_equal(col, col_o) and np.array_equal(wd, w_o)


@pytest.mark.parametrize("weighted,text", [(True, False), (False, False), (True, True), (False, True)])
def test_file_formats_load_like_the_arrays(engine, karate, tmp_path, weighted, text):
    """Binary Edge<real_t> (12-byte records) / Edge<EmptyData> (8-byte) files as
    write_graph leaves them (include/storage.hpp:57-66), and "src dst[ w]" text
    edge lists, give the CSR of the same edges passed as arrays."""
    src, dst, w = karate
    w = w if weighted else None
    path = tmp_path / ("g.txt" if text else "g.bin")
    if text:
        with open(path, "w") as f:
            for i in range(len(src)):
                f.write(f"{src[i]} {dst[i]}" + (f" {float(w[i]):.9g}" if weighted else "") + "\n")
    else:
        if weighted:
            rec = np.zeros(len(src), dtype=[("s", "<u4"), ("d", "<u4"), ("w", "<f4")])
            rec["w"] = w
        else:
            rec = np.zeros(len(src), dtype=[("s", "<u4"), ("d", "<u4")])
        rec["s"] = src
        rec["d"] = dst
        rec.tofile(path)
    print(f"[gpu] karate from a {'text' if text else 'binary'} file, weighted={weighted}", flush=True)
    engine.load_graph(34, str(path), text=text, weighted=weighted, make_undirected=False, sort_adj=True)
    rp, col, wd, _ = engine.download(weights=weighted)
    rp_o, col_o, w_o = O.csr_from_edges(34, src, dst, w, undirected=False, sort_adj=True)
    assert np.array_equal(rp, rp_o) and np.array_equal(col, col_o)
    if weighted:
        assert np.array_equal(wd.view(np.uint32), w_o.view(np.uint32))


def test_edges_csr_rejects_bad_vertex(engine):
    with pytest.raises(kk.KkwError):
        engine.graph_from_edges(4, np.array([0, 1], np.uint32), np.array([1, 9], np.uint32))


def test_failed_build_leaves_no_graph(engine):
    """A build that fails part-way (a bad vertex id found on the device)
    leaves no half-built graph: the engine reports that none is loaded."""
    engine.graph_rmat(8, 8, seed=1, weighted=True)
    with pytest.raises(kk.KkwError):
        engine.graph_from_edges(4, np.array([0, 1], np.uint32), np.array([1, 9], np.uint32))
    with pytest.raises(kk.KkwError) as e:
        engine.info()
    assert e.value.code == kk.KKW_E_NOGRAPH


def test_bad_edge_types_keep_the_previous_ones(engine):
    """Edge types of 8 or more are refused (one bit each in a u8 mask); the
    previous types stay, and a metapath walk gives the same paths as before."""
    load_rmat(engine, 10)
    engine.type_by_dst(5, 3)
    job = kk.metapath_job(2000, 12, [[0b00111, 0b11000], [0b11111]], True, seed=4)
    (before,) = engine.run([job])
    V, E = engine.info()
    bad = np.full(E, 9, np.uint8)
    with pytest.raises(kk.KkwError) as e:
        engine.set_edge_types(bad)
    assert e.value.code == kk.KKW_E_INVALID
    (after,) = engine.run([job])
    assert np.array_equal(before.lengths, after.lengths)
    mask = np.arange(before.vertices.shape[1])[None, :] < before.lengths[:, None]
    assert np.all((before.vertices == after.vertices) | ~mask)


@pytest.mark.parametrize("scale", [10, 14])
def test_alias_matches_oracle(engine, scale):
    load_rmat(engine, scale)
    engine.type_by_dst(5, 3)
    engine.prepare()
    p, prim, al, meta, vs = engine.alias_download()
    rp, col, w, et = engine.download(etype=True)
    p_o, prim_o, al_o, meta_o, vs_o = O.alias_build(rp, col, w, et)
    assert np.array_equal(p.view(np.uint32), p_o.view(np.uint32))
    assert np.array_equal(prim, prim_o)
    assert np.array_equal(al, al_o)
    assert np.array_equal(meta, meta_o)
    assert np.array_equal(vs.view(np.uint32), vs_o.view(np.uint32))


def _hub_graph(seed):
    """CSR whose long lists cover the wave-per-list alias build and its
    block-per-list form for lists of more than 4096 entries: lists just below,
    at and above 1024 and 4096 entries and up to 300k, with weight shapes that
    drive the small/large queues differently (uniform, heavy-tailed, all equal,
    one dominant weight, one tiny weight)."""
    rng = np.random.default_rng(seed)
    degs = [3, 1023, 1024, 1025, 4096, 70000, 300000, 5000, 5000, 5000, 2048, 17, 4097, 2049]
    V = len(degs) + 64
    degs = degs + [0] * (V - len(degs))
    rp = np.zeros(V + 1, np.uint64)
    rp[1:] = np.cumsum(degs)
    E = int(rp[-1])
    col = rng.integers(0, V, E).astype(np.uint32)
    w = (1.0 + 4.0 * rng.random(E)).astype(np.float32)
    def seg(i):
        return slice(int(rp[i]), int(rp[i + 1]))
    w[seg(5)] = (rng.exponential(1.0, degs[5]) ** 3 + 1e-3).astype(np.float32)
    w[seg(7)] = 2.5
    w[seg(8)] = 1.0
    w[int(rp[8])] = 1e4
    w[seg(9)] = 1.0
    w[int(rp[9]) + 17] = 1e-6
    w[seg(10)] = np.where(rng.random(degs[10]) < 0.05, 50.0, 0.5).astype(np.float32)
    return rp, col, w


@pytest.mark.parametrize("seed", [1, 2])
def test_alias_matches_oracle_long_lists(engine, seed):
    rp, col, w = _hub_graph(seed)
    print(f"[gpu] alias on long lists, seed {seed}, E={int(rp[-1])}", flush=True)
    engine.graph_from_csr(rp, col, w)
    et = np.random.default_rng(seed + 7).integers(0, 5, len(col)).astype(np.uint8)
    engine.set_edge_types(et)
    engine.prepare()
    p, prim, al, meta, vs = engine.alias_download()
    p_o, prim_o, al_o, meta_o, vs_o = O.alias_build(rp, col, w, et)
    assert np.array_equal(vs.view(np.uint32), vs_o.view(np.uint32))
    assert np.array_equal(p.view(np.uint32), p_o.view(np.uint32))
    assert np.array_equal(prim, prim_o)
    assert np.array_equal(al, al_o)
    assert np.array_equal(meta, meta_o)


def _integer_weight_graph(seed):
    """Undirected graph with integer weights rand() % 5 + 1 as the reference's
    test generator draws them (src/tests/generator_helper.hpp:65-76): ties with
    a list's average weight are common.  Hubs with 40-700 neighbours take the
    wave-per-list alias build; the rest the lane-per-list one."""
    rng = np.random.default_rng(seed)
    V = 3000
    pairs = [rng.integers(0, V, (12000, 2))]
    for h in range(12):
        k = int(rng.integers(40, 700))
        pairs.append(np.stack([np.full(k, h), rng.integers(0, V, k)], 1))
    e = np.concatenate(pairs).astype(np.int64)
    e = e[e[:, 0] != e[:, 1]]
    e = np.unique(np.sort(e, 1), axis=0)
    w = (rng.integers(0, 5, len(e)) + 1).astype(np.float32)
    return V, e[:, 0].astype(np.uint32), e[:, 1].astype(np.uint32), w


@pytest.mark.parametrize("seed", [1, 2])
def test_integer_weights_alias_and_walks(engine, seed):
    """Alias tables and walks on integer weights, bit for bit against the oracle."""
    V, src, dst, w = _integer_weight_graph(seed)
    print(f"[gpu] integer weights, seed {seed}, {len(src)} undirected edges", flush=True)
    engine.graph_from_edges(V, src, dst, w, make_undirected=True, sort_adj=True)
    engine.prepare()
    rp, col, wd, _ = engine.download()
    p, prim, al, meta, vs = engine.alias_download()
    p_o, prim_o, al_o, meta_o, vs_o = O.alias_build(rp, col, wd)
    assert np.array_equal(vs.view(np.uint32), vs_o.view(np.uint32))
    assert np.array_equal(p.view(np.uint32), p_o.view(np.uint32))
    assert np.array_equal(prim, prim_o)
    assert np.array_equal(al, al_o)
    # ties decide buckets here: some list has a weight equal to its average
    ave = np.repeat(vs_o / np.maximum(np.diff(rp.astype(np.int64)), 1), np.diff(rp.astype(np.int64)))
    assert np.any(wd == ave.astype(np.float32))
    g = O.Graph(rp, col, wd)
    for mode in MODES:
        for job in (dict(kind=kk.DEEPWALK, walker_num=6000, walk_length=40, seed=60 + seed),
                    dict(kind=kk.NODE2VEC, walker_num=6000, walk_length=40, p=2.0, q=0.5, seed=70 + seed)):
            gpu, op, ol, ot = _walk_both(engine, g, mode, **dict(job))
            assert_paths_equal(gpu, op, ol, ot, f"integer weights kind {job['kind']} mode {mode}")


def _walk_both(engine, g, mode, **job):
    kind = job.pop("kind")
    n = job.pop("walker_num")
    L = job.pop("walk_length", 0)
    weighted = job.pop("use_weights", True)
    kj = kk.Job(kind, n, L, use_weights=weighted, **job)
    (gpu,) = run_jobs(engine, [kj], mode)
    okw = dict(job)
    okw.pop("schemes", None)
    o_paths, o_lens, o_tr, st = O.walk(g, kind, n, L, weighted, schemes=job.get("schemes"), **okw)
    return gpu, o_paths, o_lens, o_tr


WALK_CASES = {
    "deepwalk_w": dict(kind=kk.DEEPWALK, walker_num=3000, walk_length=40, seed=11),
    "deepwalk_u": dict(kind=kk.DEEPWALK, walker_num=3000, walk_length=40, seed=12, use_weights=False),
    "ppr_w": dict(kind=kk.PPR, walker_num=3000, terminate_prob=1 / 80, seed=13),
    "ppr_list_u": dict(kind=kk.PPR, walker_num=2000, terminate_prob=0.15, seed=14, use_weights=False,
                       start_mode=kk.START_LIST, starts=np.array([5, 77, 300], np.uint32)),
    "n2v_w": dict(kind=kk.NODE2VEC, walker_num=3000, walk_length=40, p=2.0, q=0.5, seed=15),
    "n2v_u": dict(kind=kk.NODE2VEC, walker_num=3000, walk_length=40, p=2.0, q=0.5, seed=16, use_weights=False),
    "n2v_outlier_w": dict(kind=kk.NODE2VEC, walker_num=2000, walk_length=30, p=0.25, q=2.0, seed=17),
    "n2v_outlier_u": dict(kind=kk.NODE2VEC, walker_num=2000, walk_length=30, p=0.5, q=4.0, seed=18,
                          use_weights=False),
    "n2v_q_small": dict(kind=kk.NODE2VEC, walker_num=2000, walk_length=30, p=1.0, q=0.25, seed=19),
    "n2v_uniform_start": dict(kind=kk.NODE2VEC, walker_num=2000, walk_length=20, p=4.0, q=1.0, seed=20,
                              start_mode=kk.START_UNIFORM),
    "deepwalk_epoch": dict(kind=kk.DEEPWALK, walker_num=1500, walk_length=25, seed=21, walker_begin=1500),
    # L = 80: persistent rows of three whole 128-byte lines (line-buffered write-out)
    "deepwalk_w_l80": dict(kind=kk.DEEPWALK, walker_num=3000, walk_length=80, seed=24),
    "n2v_u_l80": dict(kind=kk.NODE2VEC, walker_num=3000, walk_length=80, p=0.5, q=2.0, seed=25, use_weights=False),
    # L = 63: full walks end on the last entry of a line (no partial line to flush)
    "deepwalk_w_l63": dict(kind=kk.DEEPWALK, walker_num=2000, walk_length=63, seed=26),
}


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("case", sorted(WALK_CASES))
def test_walk_matches_oracle(engine, rmat10, case, mode):
    load_rmat(engine, 10)
    print(f"[gpu] walk {case} mode={mode}", flush=True)
    gpu, op, ol, ot = _walk_both(engine, rmat10, mode, **dict(WALK_CASES[case]))
    assert_paths_equal(gpu, op, ol, ot, case)


SCHEMES = [[0b00011, 0b11100], [0b10101, 0b01010, 0b11111], [0b00001, 0b00010, 0b00100, 0b01000]]


@pytest.mark.parametrize("weighted", [True, False])
@pytest.mark.parametrize("mode", [kk.MODE_PLAIN, kk.MODE_PERSISTENT], ids=["plain", "persistent"])
def test_metapath_matches_oracle(engine, weighted, mode):
    load_rmat(engine, 10)
    engine.type_by_dst(5, 9)
    rp, col, w, et = engine.download(etype=True)
    g = O.Graph(rp, col, w, et)
    job = dict(kind=kk.METAPATH, walker_num=3000, walk_length=30, seed=23, use_weights=weighted, schemes=SCHEMES)
    print(f"[gpu] metapath weighted={weighted} mode={mode}", flush=True)
    gpu, op, ol, ot = _walk_both(engine, g, mode, **job)
    assert_paths_equal(gpu, op, ol, ot, "metapath")


def test_node2vec_rmat14_config(engine):
    """node2vec p=2 q=0.5 weighted, R-MAT scale 14, one walker per vertex, L=80, seed 1."""
    load_rmat(engine, 14)
    rp, col, w = O.rmat_graph(14, 16, seed=1, weighted=True)
    g = O.Graph(rp, col, w)
    op, ol, ot, st = O.walk(g, O.NODE2VEC, 16384, 80, True, p=2.0, q=0.5, seed=1)
    for mode in MODES:
        (gpu,) = run_jobs(engine, [kk.node2vec_job(16384, 80, 2.0, 0.5, True, seed=1)], mode)
        assert_paths_equal(gpu, op, ol, ot, f"rmat14 mode {mode}")
        assert gpu.steps == st.steps


def node2vec_class_chi2(P, L, rp, col, w, p, q, min_obs=40, max_states=4000):
    """Second-order law of src/tests/test_node2vec.cpp:45-79 by class: for a
    sample of (previous a, current b) states seen at least min_obs times, the
    next step returns to a, goes to a neighbour of a, or goes farther, with
    masses sum w(b,x) over each class times 1/p, 1, 1/q.  Pooled chi-square
    over the states; returns (states, statistic, dof, p-value)."""
    from scipy import stats
    P = P.astype(np.int64)
    L = L.astype(np.int64)
    keys = []
    for s in range(1, P.shape[1] - 1):
        ok = L > s + 1
        if not ok.any():
            break
        keys.append((P[ok, s - 1] << 28) | (P[ok, s] << 14) | P[ok, s + 1])
    uniq, cnt = np.unique(np.concatenate(keys), return_counts=True)
    su, first, scnt = np.unique(uniq >> 14, return_index=True, return_counts=True)
    tot = np.add.reduceat(cnt, first)
    idx = np.nonzero(tot >= min_obs)[0]
    if len(idx) > max_states:
        idx = np.random.default_rng(0).choice(idx, max_states, replace=False)
    nbr = {}

    def nb(v):
        if v not in nbr:
            nbr[v] = set(col[rp[v]:rp[v + 1]].tolist())
        return nbr[v]

    stat, dof = 0.0, 0
    for i in idx.tolist():
        a, b = int(su[i]) >> 14, int(su[i]) & 0x3FFF
        f0, k = int(first[i]), int(scnt[i])
        na = nb(a)
        obs = [0, 0, 0]
        for x, c in zip((uniq[f0:f0 + k] & 0x3FFF).tolist(), cnt[f0:f0 + k].tolist()):
            obs[0 if x == a else (1 if x in na else 2)] += c
        mass = [0.0, 0.0, 0.0]
        for e in range(int(rp[b]), int(rp[b + 1])):
            x = int(col[e])
            j = 0 if x == a else (1 if x in na else 2)
            mass[j] += float(w[e]) * (1 / p, 1.0, 1 / q)[j]
        S, N = sum(mass), sum(obs)
        cells = [(obs[j], N * mass[j] / S) for j in range(3) if mass[j] > 0]
        small_o = sum(o for o, e in cells if e < 5)
        small_e = sum(e for o, e in cells if e < 5)
        kk_ = 0
        for o, e in cells:
            if e >= 5:
                stat += (o - e) ** 2 / e
                kk_ += 1
        if small_e >= 5:
            stat += (small_o - small_e) ** 2 / small_e
            kk_ += 1
        dof += max(kk_ - 1, 0)
    return len(idx), stat, dof, stats.chi2.sf(stat, dof)


def test_node2vec_rmat14_distribution(engine):
    """On the config-1 graph the GPU's node2vec paths follow the second-order
    closed form directly (class test above, 4000 states); the same check with
    p and q exchanged in the expected law rejects them.  (Bit-exact agreement
    with the oracle, whose full closed forms are checked on small graphs,
    implies the first; this checks it at this size on the device's paths.)"""
    load_rmat(engine, 14)
    rp, col, w, _ = engine.download()
    print("[gpu] node2vec distribution on R-MAT-14, 262144 walkers", flush=True)
    (gpu,) = engine.run([kk.node2vec_job(16384 * 16, 80, 2.0, 0.5, True, seed=7)], mode=kk.MODE_PERSISTENT)
    n, stat, dof, pv = node2vec_class_chi2(gpu.vertices, gpu.lengths, rp, col, w, 2.0, 0.5)
    print(f"[gpu] {n} states, chi2 {stat:.0f} on {dof} dof, p={pv:.3g}", flush=True)
    assert n >= 1000 and dof > 1000 and pv > 1e-4
    _, _, _, pv_wrong = node2vec_class_chi2(gpu.vertices, gpu.lengths, rp, col, w, 0.5, 2.0)
    assert pv_wrong < 1e-6


def test_mixed_batch_equals_single_jobs(engine):
    """All kinds in one persistent launch (the mixed-kind kernel) give each job
    the oracle's paths, and the same paths as the job alone in a plain launch."""
    from helpers import assert_rows_match_oracle, oracle_walk
    load_rmat(engine, 10)
    engine.type_by_dst(5, 9)
    rp, col, w, et = engine.download(etype=True)
    g = O.Graph(rp, col, w, et)
    jobs = [
        kk.deepwalk_job(700, 33, True, seed=31),
        kk.ppr_job(900, 0.05, True, seed=32, start_vertices=np.array([1, 2, 3], np.uint32)),
        kk.node2vec_job(1100, 25, 0.5, 2.0, True, seed=33),
        kk.metapath_job(500, 20, SCHEMES, True, seed=34),
        kk.node2vec_job(300, 15, 2.0, 0.5, False, seed=35),
        kk.deepwalk_job(1, 10, False, seed=36),
        kk.deepwalk_job(0, 10, False, seed=37),
        kk.deepwalk_job(800, 63, True, seed=38),   # 64-entry rows: half-line write-out
    ]
    mixed = engine.run(jobs, mode=kk.MODE_PERSISTENT)
    for i, (j, m) in enumerate(zip(jobs, mixed)):
        op, ol, ot, _ = oracle_walk(g, j)
        assert_rows_match_oracle(m, op, ol, ot, what=f"job {i} in the mix")
        (alone,) = engine.run([j], mode=kk.MODE_PLAIN)
        assert_same_rows(alone, m)


@pytest.mark.parametrize("n_jobs", [64, 80], ids=["table_in_lds", "table_in_hbm"])
def test_mixed_batch_many_jobs(engine, n_jobs):
    """Batches up to 64 jobs read their job table from a per-block LDS copy,
    larger ones from device memory: both give every job the oracle's paths.
    Metapath jobs carry different scheme sets, so each walker's scheme masks
    (held in registers from its start) come from its own job."""
    from helpers import assert_rows_match_oracle, oracle_walk
    load_rmat(engine, 10)
    engine.type_by_dst(5, 9)
    rp, col, w, et = engine.download(etype=True)
    g = O.Graph(rp, col, w, et)
    rng = np.random.default_rng(n_jobs)
    jobs = []
    for i in range(n_jobs):
        n = int(rng.integers(20, 300))
        k = i % 4
        if k == 0:
            jobs.append(kk.deepwalk_job(n, int(rng.integers(5, 40)), bool(i & 4), seed=500 + i))
        elif k == 1:
            jobs.append(kk.ppr_job(n, 0.1, True, seed=500 + i))
        elif k == 2:
            schemes = [[int(rng.integers(1, 32)) for _ in range(int(rng.integers(1, 9)))]
                       for _ in range(int(rng.integers(1, 4)))]
            jobs.append(kk.metapath_job(n, int(rng.integers(5, 40)), schemes, True, seed=500 + i))
        else:
            jobs.append(kk.node2vec_job(n, int(rng.integers(5, 30)), 2.0, 0.5, True, seed=500 + i))
    print(f"[gpu] {n_jobs} mixed jobs on R-MAT-10 in one persistent launch", flush=True)
    mixed = engine.run(jobs, mode=kk.MODE_PERSISTENT)
    for i, (j, m) in enumerate(zip(jobs, mixed)):
        op, ol, ot, _ = oracle_walk(g, j)
        assert_rows_match_oracle(m, op, ol, ot, what=f"job {i} of {n_jobs}")


def test_graph_replay_matches(engine):
    """A batch captured in a hipGraph and replayed gives the same paths every time."""
    load_rmat(engine, 10)
    jobs = [kk.node2vec_job(2048, 30, 2.0, 0.5, True, seed=41), kk.ppr_job(1024, 0.1, True, seed=42)]
    ref = engine.run(jobs, mode=kk.MODE_PLAIN)
    b = engine.batch(jobs, kk.MODE_PERSISTENT, kk.BATCH_GRAPH)
    for rep in range(3):
        b.launch()
        b.sync()
        for i in range(2):
            assert_same_rows(b.paths(i), ref[i], f"replay {rep} job {i}")
    b.close()


def test_stats_variant_same_paths(engine):
    load_rmat(engine, 10)
    job = kk.node2vec_job(2048, 30, 2.0, 0.5, True, seed=43)
    (a,) = engine.run([job], mode=kk.MODE_PERSISTENT)
    (b,) = engine.run([job], mode=kk.MODE_PERSISTENT, stats=True)
    t = engine.timing()
    assert np.array_equal(a.lengths, b.lengths)
    assert t["steps"] == a.steps and t["trials"] >= t["steps"]


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_node2vec_trial_cap(engine, mode):
    """A return step accepted only when the dart is exactly 0 (p = 1e30 on a
    single edge: 1/p = 1e-30, darts on the 2^-24 grid): most walkers spend the
    per-step cap of 2^20 trials and end there, truncated, after the steps they
    had.  GPU and oracle agree on every row, length and flag, in both launch
    forms (the state-machine kernel writes such a walker's last line as it
    writes the last line of any walk)."""
    src, dst = np.array([0], np.uint32), np.array([1], np.uint32)
    engine.graph_from_edges(2, src, dst, np.array([2.0], np.float32), make_undirected=True, sort_adj=True)
    rp, col, w, _ = engine.download()
    g = O.Graph(rp, col, w)
    print(f"[gpu] node2vec trial cap, mode={mode}", flush=True)
    (gpu,) = run_jobs(engine, [kk.node2vec_job(16, 10, 1e30, 1.0, True, seed=5)], mode)
    op, ol, ot, _ = O.walk(g, O.NODE2VEC, 16, 10, True, p=1e30, q=1.0, seed=5)
    assert ot.sum() >= 8, "the cap was expected to end most walks"
    assert_paths_equal(gpu, op, ol, ot, "trial cap")


def test_node2vec_needs_sorted_lists(engine, karate):
    src, dst, w = (a[::-1].copy() for a in karate)   # file order descending: lists unsorted
    engine.graph_from_edges(34, src, dst, w, sort_adj=False)
    with pytest.raises(kk.KkwError):
        engine.run([kk.node2vec_job(10, 5, 2.0, 0.5)])


def test_large_node2vec_properties(engine):
    """R-MAT scale 20, node2vec p=2 q=0.5 weighted, one walker per vertex, L=80:
    every path follows CSR edges, non-isolated walkers take exactly 80 steps,
    isolated ones none; plain and persistent launches agree."""
    load_rmat(engine, 20)
    rp, col, _, _ = engine.download(weights=False)
    V = len(rp) - 1
    job = kk.node2vec_job(V, 80, 2.0, 0.5, True, seed=5)
    (pp,) = engine.run([job], mode=kk.MODE_PERSISTENT)
    deg = np.diff(rp.astype(np.int64))
    expect = np.where(deg[np.arange(V)] > 0, 81, 1)
    assert np.array_equal(pp.lengths, expect)
    # edge check on a sample of walkers: (u,v) present iff binary search finds v in u's list
    rng = np.random.default_rng(0)
    sample = rng.choice(np.nonzero(deg > 0)[0], 20000, replace=False)
    u = pp.vertices[sample, :80].reshape(-1).astype(np.int64)
    v = pp.vertices[sample, 1:81].reshape(-1).astype(np.int64)
    key = (np.repeat(np.arange(V, dtype=np.int64), deg) << 32) | col.astype(np.int64)
    probe = (u << 32) | v
    pos = np.searchsorted(key, probe)
    assert np.all(key[np.minimum(pos, len(key) - 1)] == probe)
    (pl,) = engine.run([job], mode=kk.MODE_PLAIN)
    assert np.array_equal(pl.lengths, pp.lengths)
    assert np.array_equal(pl.vertices[sample, :81], pp.vertices[sample, :81])
    # 2^20 walkers: the trial-loop kernel by default; the state machine too
    (ps,) = run_jobs(engine, [job], PERSISTENT_SM)
    assert_same_rows(ps, pp, "node2vec R-MAT-20, state machine")


def _edge_keys(rp, col):
    """Sorted (u << 32 | v) keys of the CSR's edges (lists sorted by neighbour)."""
    deg = np.diff(rp.astype(np.int64))
    return (np.repeat(np.arange(len(rp) - 1, dtype=np.int64), deg) << 32) | col.astype(np.int64)


def _edges_exist(key, u, v):
    """(u, v) is an edge for every pair: binary search in the sorted keys."""
    probe = (np.asarray(u, np.int64) << 32) | np.asarray(v, np.int64)
    pos = np.searchsorted(key, probe)
    return bool(np.all(key[np.minimum(pos, len(key) - 1)] == probe))


def test_deepwalk_rmat22_properties(engine):
    """The DeepWalk workload at full size (R-MAT scale 22, 4,194,304 walkers,
    L = 80, weighted): non-isolated walkers take exactly 80 steps, isolated
    ones none; every sampled transition is an edge; the persistent and plain
    launches agree on every row."""
    load_rmat(engine, 22)
    rp, col, _, _ = engine.download(weights=False)
    V = len(rp) - 1
    deg = np.diff(rp.astype(np.int64))
    job = kk.deepwalk_job(V, 80, True, seed=1)
    (pp,) = engine.run([job], mode=kk.MODE_PERSISTENT)
    assert np.array_equal(pp.lengths, np.where(deg > 0, 81, 1))
    sample = np.random.default_rng(1).choice(np.nonzero(deg > 0)[0], 20000, replace=False)
    assert _edges_exist(_edge_keys(rp, col), pp.vertices[sample, :80].reshape(-1),
                        pp.vertices[sample, 1:81].reshape(-1))
    (pl,) = engine.run([job], mode=kk.MODE_PLAIN)
    assert_same_rows(pl, pp, "deepwalk R-MAT-22")


def test_mixed_jobs_rmat23_properties(engine):
    """The batched workload at full size (R-MAT scale 23, 5 vertex types, the
    64 mixed PPR / DeepWalk / metapath jobs of the benchmark): one persistent
    launch gives every job the rows of its own plain launch, and sampled
    metapath paths follow CSR edges whose types their schemes allow."""
    import bench
    engine.graph_rmat(23, 16, 0.57, 0.19, 0.19, seed=2, weighted=True)
    engine.type_by_dst(5, 3)
    rp, col, _, et = engine.download(weights=False, etype=True)
    vtype = O.vertex_types(len(rp) - 1, 5, 3)
    assert np.array_equal(et, vtype[col])
    from helpers import assert_rows_match_oracle, oracle_walk
    jobs, kinds = bench.mixed_jobs(kk)
    print(f"[gpu] 64 mixed jobs on R-MAT-23, {sum(j.walker_num for j in jobs)} walkers", flush=True)
    mixed = engine.run(jobs, mode=kk.MODE_PERSISTENT)
    _, _, w, _ = engine.download(weights=True)
    g = O.Graph(rp, col, w, et)
    rng = np.random.default_rng(2)
    us, vs = [], []
    for j, (job, kind, m) in enumerate(zip(jobs, kinds, mixed)):
        # every job's rows of a strided sample of about 2000 walkers against the oracle
        stride = max(1, job.walker_num // 2000)
        n = (job.walker_num + stride - 1) // stride
        op, ol, ot, _ = oracle_walk(g, job, walker_num=n, walker_stride=stride)
        assert_rows_match_oracle(m, op, ol, ot, np.arange(n) * stride, f"job {j} ({kind}) in the mix")
        (alone,) = engine.run([job], mode=kk.MODE_PLAIN)
        assert_same_rows(alone, m, f"job {j} ({kind})")
        if kind != "metapath":
            continue
        long_rows = np.nonzero(m.lengths > 1)[0]
        rows = rng.choice(long_rows, min(200, len(long_rows)), replace=False)
        for r in rows:
            path = m.vertices[r, :m.lengths[r]].astype(np.int64)
            us.append(path[:-1])
            vs.append(path[1:])
            # the walker's scheme (its INIT draw) allows the type of every step's edge
            wid = job.walker_begin + int(r)
            d = O.philox([wid, 0, 0, 2], [job.seed & 0xFFFFFFFF, job.seed >> 32])
            scheme = job.schemes[O.bounded(int(d[0]), int(d[1]), len(job.schemes))]
            for s in range(len(path) - 1):
                t_e = int(vtype[path[s + 1]])   # edge type = type of its destination
                assert (scheme[s % len(scheme)] >> t_e) & 1, (j, r, s)
    assert _edges_exist(_edge_keys(rp, col), np.concatenate(us), np.concatenate(vs))


@pytest.mark.parametrize("p,q", [(2.0, 0.5), (0.5, 2.0), (1.0, 0.25)])
@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_node2vec_directed_graph(engine, p, q, mode):
    """An asymmetric graph: the adjacency test must search the previous vertex's
    list (no shortcut through the candidate's list); where the return edge is
    missing the outlier appendix is empty on both sides."""
    rng = np.random.default_rng(7)
    V, n = 300, 4000
    src = rng.integers(0, V, n).astype(np.uint32)
    dst = rng.integers(0, V, n).astype(np.uint32)
    w = (1 + 4 * rng.random(n)).astype(np.float32)
    engine.graph_from_edges(V, src, dst, w, make_undirected=False, sort_adj=True)
    rp, col, wd, _ = engine.download()
    g = O.Graph(rp, col, wd)
    print(f"[gpu] directed node2vec p={p} q={q} mode={mode}", flush=True)
    (gpu,) = run_jobs(engine, [kk.node2vec_job(2000, 30, p, q, True, seed=9)], mode)
    op, ol, ot, st = O.walk(g, O.NODE2VEC, 2000, 30, True, p=p, q=q, seed=9)
    assert_paths_equal(gpu, op, ol, ot, "directed")


def test_captured_batch_follows_graph_changes(engine):
    """A hipGraph-captured batch re-captures after the neighbour hash sets are
    dropped or rebuilt (same paths), and refuses to run on a replaced graph."""
    load_rmat(engine, 10)
    job = kk.node2vec_job(2048, 20, 2.0, 0.5, True, seed=51)
    (ref,) = engine.run([job], mode=kk.MODE_PLAIN)
    b = engine.batch([job], kk.MODE_PERSISTENT, kk.BATCH_GRAPH)
    for enable in (True, False, True):
        engine.adj_index(enable)
        b.launch()
        b.sync()
        assert_same_rows(b.paths(0), ref)
    load_rmat(engine, 9)
    with pytest.raises(kk.KkwError):
        b.launch()
    b.close()


@pytest.mark.parametrize("index", [True, False])
def test_node2vec_with_and_without_hash_sets(engine, index):
    """The adjacency test answers the same through the neighbour hash sets and
    through the binary search (both launch forms, weighted and unweighted)."""
    load_rmat(engine, 12)
    engine.adj_index(index)
    rp, col, w = O.rmat_graph(12, 16, seed=1, weighted=True)
    g = O.Graph(rp, col, w)
    for weighted in (True, False):
        op, ol, ot, st = O.walk(g, O.NODE2VEC, 4096, 40, weighted, p=2.0, q=0.5, seed=61)
        for mode in MODES:
            (gpu,) = run_jobs(engine, [kk.node2vec_job(4096, 40, 2.0, 0.5, weighted, seed=61)], mode)
            assert_paths_equal(gpu, op, ol, ot, f"index={index} weighted={weighted} mode={mode}")
    engine.adj_index(True)


ADJ_SEG = 16384   # slots of one LDS segment of a long list's hash set


def adj_home(keys, size):
    """Home slot of each key in a hash set of `size` slots (the layout the
    library documents in include/kkw_program.h)."""
    k = np.asarray(keys, np.uint64)
    return (((k * np.uint64(0x9E3779B1)) & np.uint64(0xFFFFFFFF)) * np.uint64(size)) >> np.uint64(32)


def segment_overflows(rp, col, min_deg=4096):
    """Keys of lists of min_deg or more entries whose linear probe, run inside
    their own ADJ_SEG-slot segment, passes the segment's end."""
    rp = rp.astype(np.int64)
    total = 0
    for v in np.nonzero(np.diff(rp) >= min_deg)[0]:
        keys = col[rp[v]:rp[v + 1]]
        size = 4 * len(keys) + 4
        h = adj_home(keys, size).astype(np.int64)
        for s0 in range(0, size, ADJ_SEG):
            w = min(ADJ_SEG, size - s0)
            pos = -1
            for x in np.sort(h[(h >= s0) & (h < s0 + w)] - s0):
                pos = max(pos + 1, int(x))
                total += pos >= w
    return total


def hub_graph(V=200000, extra=300000, seed=5):
    """Undirected weighted graph whose hubs have lists far longer than one LDS
    table: vertex 0 joined to every vertex, vertex 1 to 20,000 more (all
    candidates hashing into the last 8 slots of the first segment of its table
    among them, so keys run past a segment's end), plus random edges; no
    duplicates, no self loops, lists sorted."""
    rng = np.random.default_rng(seed)
    cand = np.arange(2, V)
    home = adj_home(cand, 4 * 20001 + 4)   # vertex 1's list: 20,000 picks and vertex 0
    hot = cand[(home >= ADJ_SEG - 8) & (home < ADJ_SEG)]   # about 20 keys for 8 slots
    rest = rng.permutation(np.setdiff1d(cand, hot))[:20000 - len(hot)]
    picks = np.concatenate([hot, rest])
    u = np.concatenate([np.zeros(V - 1, np.int64), np.ones(len(picks), np.int64), rng.integers(2, V, extra)])
    v = np.concatenate([np.arange(1, V), picks, rng.integers(2, V, extra)])
    keep = u != v
    key = np.unique(np.minimum(u, v)[keep] * V + np.maximum(u, v)[keep])
    src, dst = (key // V).astype(np.uint32), (key % V).astype(np.uint32)
    w = rng.uniform(0.1, 2.0, len(src)).astype(np.float32)
    return O.csr_from_edges(V, src, dst, w, undirected=True, sort_adj=True)


@pytest.mark.parametrize("index", ["segments", "fallback", "off"])
def test_node2vec_hash_sets_of_hub_lists(engine, index, monkeypatch):
    """Adjacency tests against hub lists of 199,999 and 20,001 entries (their
    hash sets span many 64 KiB LDS segments) answer as the binary search does:
    paths bit-exact against the oracle, both launch forms.  "fallback" leaves
    no room for keys that run past a segment, so every key of the long lists
    is inserted again with global atomics."""
    rp, col, w = hub_graph()
    deg = np.diff(rp.astype(np.int64))
    assert deg[0] == len(deg) - 1 and deg[1] == 20001
    assert segment_overflows(rp, col) >= 8   # about 20 keys homed in 8 slots before a segment's end
    engine.graph_from_csr(rp, col, w)
    if index == "fallback":
        monkeypatch.setenv("KKW_ADJ_OVERFLOW_CAP", "0")   # read when the hash sets are built
    engine.adj_index(index != "off")
    monkeypatch.delenv("KKW_ADJ_OVERFLOW_CAP", raising=False)
    g = O.Graph(rp, col, w)
    n = 30000
    for weighted in (True, False):
        op, ol, ot, st = O.walk(g, O.NODE2VEC, n, 20, weighted, p=2.0, q=0.5, seed=71)
        for mode in MODES:
            print(f"[gpu] hub lists index={index} weighted={weighted} mode={mode}", flush=True)
            (gpu,) = run_jobs(engine, [kk.node2vec_job(n, 20, 2.0, 0.5, weighted, seed=71)], mode)
            assert_paths_equal(gpu, op, ol, ot, f"hubs index={index} weighted={weighted} mode={mode}")
    engine.adj_index(True)


TORCH_HANDOFF = r"""
import sys
import numpy as np
import torch
torch.cuda.init()                     # torch's HIP runtime first: libkkw binds to it
sys.path.insert(0, sys.argv[1])
import knightking_amd as kk
eng = kk.Engine()
eng.graph_rmat(10, 16, seed=1, weighted=True)
b = eng.batch([kk.node2vec_job(3000, 80, 2.0, 0.5, True, seed=61)], kk.MODE_PERSISTENT)
b.launch()
b.sync()
host = b.paths(0)
tv, tl = b.paths_torch(0)
assert tv.is_cuda and tv.dtype == torch.int32 and tuple(tv.shape) == host.vertices.shape
assert np.array_equal((tl & 0x7fffffff).cpu().numpy().view(np.uint32), host.lengths)
hmask = np.arange(host.vertices.shape[1])[None, :] < host.lengths[:, None]
assert np.array_equal(tv.cpu().numpy().view(np.uint32)[hmask], host.vertices[hmask])
mask = torch.arange(tv.shape[1], device=tv.device)[None, :] < (tl & 0x7fffffff)[:, None]
visits = torch.bincount(tv[mask].long(), minlength=1024).cpu().numpy()
assert np.array_equal(visits, np.bincount(host.vertices[hmask].astype(np.int64), minlength=1024))
del tv, tl
b.close()
eng.close()
print("handoff ok")
"""


def test_device_paths_as_torch_tensors():
    """Batch paths handed to torch where the kernel wrote them (no copy): the
    same rows as the host copy, usable by torch ops on the GPU (visit counts).
    Its own process, with torch's GPU initialised before the engine."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    print("[gpu] torch handoff in a child process", flush=True)
    r = subprocess.run([sys.executable, "-c", TORCH_HANDOFF, root], capture_output=True, text=True, timeout=100)
    assert r.returncode == 0 and "handoff ok" in r.stdout, r.stdout[-2000:] + r.stderr[-2000:]
