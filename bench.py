#!/usr/bin/env python3
"""Benchmark: walker steps/s of node2vec (p=2, q=0.5, weighted) on an undirected
R-MAT scale-24 graph on one MI355X, persistent work-queue kernel (the value)
beside the plain one-job-per-launch kernel, plus p50/p99 latency of PPR
queries through the serving loop.

One "step" of this benchmark is one pass of the hot path over one batch: every
walker of the job (one per vertex, length 80) walked to the end, graph and
alias tables resident in HBM.  The CPU baseline is the oracle (the C
restatement) on the box's host cores, on a strided sample of the same walkers;
the sampled paths are also compared with the GPU's, bit for bit.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, published
LINE = 64                   # bytes per random access in the reference's per-step model
L2_LINE = 128               # bytes one L2 miss fetches on MI355X (TCC_EA0_RDREQ are 128 B)
CALIB_BYTES = 1 << 30       # streaming-copy calibration size (read once, written once)
GATHER_BYTES = 16 << 30     # random-gather ceiling: buffer about the size of the R-MAT-24 alias tables


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def host_cpus():
    """The host's CPU model and counts, and the threads the CPU baseline uses:
    every CPU this process may run on (affinity mask, cgroup CPU quota), capped
    by OMP_NUM_THREADS where the host sets it (the CPU share of one GPU on a
    shared box)."""
    info = {"cpu_model": None, "nproc": os.cpu_count(), "affinity_cpus": None, "cgroup_cpus": None,
            "omp_num_threads": None}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["cpu_model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                info["cgroup_cpus"] = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        info["omp_num_threads"] = int(omp)
    caps = [info[k] for k in ("nproc", "affinity_cpus", "cgroup_cpus", "omp_num_threads") if info[k]]
    info["threads"] = max(1, min(caps)) if caps else 1
    return info


def source_digest():
    """Digest of the sources that decide the walk kernel's memory traffic (the
    kernels, the table layouts and their construction)."""
    import hashlib
    h = hashlib.sha256()
    for name in ("walk.hip", "graph.hip", "kkw_dev.h", "kkw_rng.h"):
        with open(os.path.join(ROOT, "knightking_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def reference_model_bytes(t, weighted):
    """Algorithmic bytes of one launch under the per-step model of KnightKing's
    node2vec loop: per step 64 B (row_ptr pair) + 4 B (path entry); per trial
    64 B alias bucket + 64 B neighbour entry (weighted) or 64 B neighbour entry
    (unweighted); per adjacency query the reference would send, 64 B times
    max(1, ceil(log2(deg(prev)+1)) - 3) binary-search lines.  T, f and L_b are
    the counts measured by the instrumented run of the same launch."""
    per_trial = 2 * LINE if weighted else LINE
    return (LINE + 4) * t["steps"] + per_trial * t["trials"] + LINE * t["ref_query_lines"]


def path_bytes(t, walkers, stride=84):
    """Path rows (whole 16-byte groups actually written) plus lengths."""
    groups = (t["steps"] + walkers + 3 * walkers) // 4   # upper bound: one partial group per walker
    return min(16 * groups, 4 * stride * walkers) + 4 * walkers


def kernel_bytes(t, walkers, line):
    """Bytes this engine's kernel must move: one `line` per data-dependent
    random access (row_ptr pairs read at walker starts, alias buckets -- a
    32-byte bucket lies in one line -- and each distinct block of an adjacency
    probe), plus the path and length bytes written."""
    return line * (t["row_loads"] + t["trials"] + t["search_lines"]) + path_bytes(t, walkers)


def time_batch(b, steps, warmup):
    for _ in range(warmup):
        b.launch()
        b.sync()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        b.launch()
        kms.append(b.sync())
    wall = time.perf_counter() - t0
    return wall, kms


def cpu_baseline(eng, job_args, gpu_batch, V, seconds=20.0):
    """Oracle on host cores over a strided sample of the walkers; returns
    (steps/s, threads, sample description, parity summary)."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    rp, col, w, _ = eng.download(weights=True)
    p, prim, al, meta, vs = eng.alias_download()
    g = O.Graph(rp, col, w, None, (p, prim, al, meta, vs))
    log(f"cpu baseline: graph and alias tables copied to host in {time.perf_counter() - t0:.1f}s")
    threads = host_cpus()["threads"]
    # calibrate on a small sample, then size the timed sample to ~`seconds`
    stride = 4096
    n = V // stride
    t0 = time.perf_counter()
    _, _, _, st = O.walk(g, O.NODE2VEC, n, 80, True, p=2.0, q=0.5, seed=job_args["seed"], threads=threads,
                         walker_stride=stride)
    dt = max(time.perf_counter() - t0, 1e-3)
    per_walker = dt / max(n, 1)
    n = int(min(V, max(1000, seconds / per_walker)))
    stride = max(1, V // n)
    n = V // stride
    t0 = time.perf_counter()
    paths, lens, tr, st = O.walk(g, O.NODE2VEC, n, 80, True, p=2.0, q=0.5, seed=job_args["seed"],
                                 threads=threads, walker_stride=stride)
    dt = time.perf_counter() - t0
    rate = st.steps / dt
    # parity of the sample against the GPU's rows
    gp = gpu_batch.paths(0)
    rows = np.arange(n, dtype=np.int64) * stride
    glens = gp.lengths[rows]
    same_len = np.array_equal(glens, lens)
    mask = np.arange(81)[None, :] < lens[:, None]
    same = same_len and bool(np.all((gp.vertices[rows][:, :81] == paths) | ~mask))
    sample = (f"node2vec p=2 q=0.5 weighted on the same R-MAT-24 graph, walkers with id = k*{stride} "
              f"(k < {n}), L=80, {st.steps} steps in {dt:.1f}s, {threads} threads")
    parity = {"walkers": int(n), "bit_exact": bool(same)}
    return rate, threads, sample, parity


def mixed_jobs(kk, n_jobs=64, seed=4, n_types=5):
    """64 jobs: kind uniform over PPR (t = 1/80), DeepWalk (L in [40,120]) and
    metapath (schemes of 3 to 5 states, each a non-empty subset of the types,
    L in [40,120]); walkers round(10^(3+3U)); every job has its own seed."""
    rng = np.random.default_rng(seed)
    jobs, kinds = [], []
    for i in range(n_jobs):
        walkers = int(round(10 ** (3 + 3 * rng.random())))
        kind = int(rng.integers(0, 3))
        L = int(rng.integers(40, 121))
        if kind == 0:
            jobs.append(kk.ppr_job(walkers, 1.0 / 80, True, seed=100 + i))
        elif kind == 1:
            jobs.append(kk.deepwalk_job(walkers, L, True, seed=100 + i))
        else:
            schemes = [[int(rng.integers(1, 1 << n_types)) for _ in range(int(rng.integers(3, 6)))]
                       for _ in range(4)]
            jobs.append(kk.metapath_job(walkers, L, schemes, True, seed=100 + i))
        kinds.append(("ppr", "deepwalk", "metapath")[kind])
    return jobs, kinds


def batched_vs_plain(kk, steps, warmup, scale=23):
    """The 64 mixed jobs in one persistent launch vs one plain launch per job."""
    eng = kk.Engine()
    try:
        eng.graph_rmat(scale, 16, 0.57, 0.19, 0.19, seed=2, weighted=True)
        eng.type_by_dst(5, 3)
        eng.prepare()
        jobs, kinds = mixed_jobs(kk)
        out = {"graph": f"R-MAT scale {scale}, 5 vertex types (edge type = type of its destination)",
               "jobs": len(jobs), "walkers": int(sum(j.walker_num for j in jobs)),
               "kinds": {k: kinds.count(k) for k in set(kinds)}}
        for name, mode, flags in (("persistent", kk.MODE_PERSISTENT, kk.BATCH_GRAPH), ("plain", kk.MODE_PLAIN, 0)):
            b = eng.batch(jobs, mode, flags)
            if mode == kk.MODE_PERSISTENT:
                out["launch_shape"] = dict(zip(("block", "blocks_per_cu", "chunk"), b.autotune(None)))
            wall, kms = time_batch(b, steps, warmup)
            total = sum(b.steps(i) for i in range(len(jobs)))
            out[name + "_steps_per_s"] = round(total * steps / wall, 1)
            out[name + "_ms"] = round(float(np.mean(kms)), 3)
            out["steps_per_launch"] = int(total)
            b.close()
        out["persistent_over_plain"] = round(out["persistent_steps_per_s"] / out["plain_steps_per_s"], 3)
        log(f"batched mixed jobs: {out}")
        return out
    finally:
        eng.close()


def deepwalk_line(kk, steps, warmup, scale=22):
    """DeepWalk with weighted alias sampling on R-MAT-22, 4,194,304 walkers, L=80."""
    eng = kk.Engine()
    try:
        eng.graph_rmat(scale, 16, 0.57, 0.19, 0.19, seed=1, weighted=True)
        eng.prepare()
        job = kk.deepwalk_job(1 << scale, 80, True, seed=1)
        out = {}
        for name, mode, flags in (("persistent", kk.MODE_PERSISTENT, kk.BATCH_GRAPH), ("plain", kk.MODE_PLAIN, 0)):
            b = eng.batch([job], mode, flags)
            wall, kms = time_batch(b, steps, warmup)
            s = b.steps(0)
            out[name + "_steps_per_s"] = round(s * steps / wall, 1)
            out["steps_per_launch"] = int(s)
            b.close()
        log(f"deepwalk R-MAT-{scale}: {out}")
        return out
    finally:
        eng.close()


def unweighted_line(kk, eng, V, steps, warmup):
    """node2vec p=2 q=0.5 without edge weights (uniform static component) on the
    resident R-MAT-24 graph, walkers = |V|, L=80: persistent vs plain."""
    job = kk.node2vec_job(V, 80, 2.0, 0.5, False, seed=1)
    out = {}
    for name, mode, flags in (("persistent", kk.MODE_PERSISTENT, kk.BATCH_GRAPH), ("plain", kk.MODE_PLAIN, 0)):
        b = eng.batch([job], mode, flags)
        if mode == kk.MODE_PERSISTENT:
            b.set_shape(64, 16, 64)
        wall, _ = time_batch(b, steps, warmup)
        s = b.steps(0)
        out[name + "_steps_per_s"] = round(s * steps / wall, 1)
        out["steps_per_launch"] = int(s)
        b.close()
    out["persistent_over_plain"] = round(out["persistent_steps_per_s"] / out["plain_steps_per_s"], 3)
    log(f"node2vec unweighted R-MAT-24: {out}")
    return out


def program_line(kk, eng, V, builtin_paths):
    """node2vec p=2 q=0.5 weighted written as a walk program
    (knightking_amd/programs/node2vec.cpp, compiled at run time), on the
    resident R-MAT-24 graph: kernel time of one run and equality with the
    built-in kernel's paths."""
    with open(os.path.join(ROOT, "knightking_amd", "programs", "node2vec.cpp")) as f:
        src = f.read()
    n = builtin_paths.vertices.shape[0]
    mask = np.arange(81)[None, :] < builtin_paths.lengths[:, None]
    out = {"adjacency": "kkw_adjacent (neighbour hash sets)",
           "tables": "the engine's buckets (ids + rows; the program reads no edge index)"}
    up, lo, ip, iq, ov = kk.node2vec_constants(2.0, 0.5)
    params = [80, up, lo, ip, iq, ov, 1.0]
    # the program as shipped, then rebuilt to draw from local-index tables
    for prefix, options, modes in (("", None, (("plain", kk.MODE_PLAIN), ("persistent", kk.MODE_PERSISTENT))),
                                   ("index_tables_", "-DKKW_EDGE_INDEX=1", (("persistent", kk.MODE_PERSISTENT),))):
        t0 = time.perf_counter()
        prog = eng.program(src, options)
        out[prefix + "compile_s"] = round(time.perf_counter() - t0, 2)
        try:
            prog.run(1024, 80, use_weights=True, seed=1, second_order=True, params=params)   # tables, warm-up
            for name, mode in modes:
                got = prog.run(V, 80, use_weights=True, seed=1, second_order=True, params=params, mode=mode)
                kms = eng.timing()["kernel_ms"]
                same = bool(np.array_equal(got.lengths, builtin_paths.lengths) and
                            np.all((got.vertices[:n, :81] == builtin_paths.vertices[:, :81]) | ~mask))
                out[prefix + name] = {"steps_per_s": round(got.steps / (kms / 1e3), 1), "kernel_ms": round(kms, 3),
                                      "same_paths_as_builtin": same}
                del got
        finally:
            prog.close()
    log(f"node2vec as a walk program: {out}")
    return out


def rmat14_line(kk, steps, warmup):
    """The reference's CPU-runnable case: node2vec p=2 q=0.5 weighted on
    R-MAT-14, 16,384 walkers (one per vertex), L=80, seed 1 -- GPU launches
    beside the oracle on the host cores, whole job, with parity."""
    from oracle import oracle as O
    eng = kk.Engine()
    try:
        eng.graph_rmat(14, 16, 0.57, 0.19, 0.19, seed=1, weighted=True)
        eng.prepare()
        V, _ = eng.info()
        job = kk.node2vec_job(V, 80, 2.0, 0.5, True, seed=1)
        out = {}
        for name, mode, flags in (("persistent", kk.MODE_PERSISTENT, kk.BATCH_GRAPH), ("plain", kk.MODE_PLAIN, 0)):
            b = eng.batch([job], mode, flags)
            wall, kms = time_batch(b, steps, warmup)
            s = b.steps(0)
            out[name + "_steps_per_s"] = round(s * steps / wall, 1)
            out[name + "_kernel_ms"] = round(float(np.mean(kms)), 4)
            out["steps_per_job"] = int(s)
            if mode == kk.MODE_PERSISTENT:
                gp = b.paths(0)
            b.close()
        rp, col, w, _ = eng.download(weights=True)
        g = O.Graph(rp, col, w)
        g.with_alias()
        threads = host_cpus()["threads"]
        reps, dt = 0, 0.0
        while dt < 2.0:
            t0 = time.perf_counter()
            paths, lens, _, st = O.walk(g, O.NODE2VEC, V, 80, True, p=2.0, q=0.5, seed=1, threads=threads)
            dt += time.perf_counter() - t0
            reps += 1
        out["cpu_oracle_steps_per_s"] = round(st.steps * reps / dt, 1)
        out["cpu_threads"] = threads
        mask = np.arange(81)[None, :] < lens[:, None]
        out["bit_exact"] = bool(np.array_equal(gp.lengths, lens) and np.all((gp.vertices[:, :81] == paths) | ~mask))
        log(f"node2vec R-MAT-14 (CPU-path case): {out}")
        return out
    finally:
        eng.close()


def serve_stream(eng, jobs, gaps):
    """Submit `jobs` at the given inter-arrival gaps while a second thread
    collects results in submission order (steps and latency; paths are freed
    unread); returns (latencies ms, steps, span s, error or None).  The
    submitter sleeps to within 0.2 ms of each arrival and spins the rest, and
    the interpreter switches threads every 0.1 ms, so the collector is never
    starved of the interpreter lock while the submitter waits."""
    n_q = len(jobs)
    tickets, lat, steps, failure = [], [], [0], []
    done = threading.Event()
    c_jobs = [j.to_c() for j in jobs]   # descriptors built before the clock starts

    def collect():
        i = 0
        try:
            while i < n_q:
                if i >= len(tickets):
                    time.sleep(0.0001)
                    continue
                s, ms = eng.wait_steps(tickets[i], 30.0)
                lat.append(ms)
                steps[0] += s
                i += 1
        except Exception as e:   # a stalled or failed query ends this measurement
            failure.append(repr(e))
        done.set()

    old_switch = sys.getswitchinterval()
    sys.setswitchinterval(1e-4)
    try:
        th = threading.Thread(target=collect)
        th.start()
        t_start = time.perf_counter()
        t_next = t_start
        for i in range(n_q):
            t_next += gaps[i]
            ahead = t_next - time.perf_counter()
            if ahead > 3e-4:
                time.sleep(ahead - 2e-4)
            while time.perf_counter() < t_next:
                pass
            tickets.append(eng.submit_c(c_jobs[i]))
        done.wait()
        th.join()
    finally:
        sys.setswitchinterval(old_switch)
    return np.array(lat), steps[0], time.perf_counter() - t_start, (failure[0] if failure else None)


def serve_session(eng, jobs, gaps, round_steps=0, groups=None, idle_poll_us=0):
    """One serving session over a query stream, with the loop's flight trace:
    returns the latency / throughput summary and the flights' occupancy
    (and p50/p99 of each named group of queries: name -> boolean mask)."""
    import tempfile
    fd, trace = tempfile.mkstemp(suffix=".csv")
    os.close(fd)
    os.environ["KKW_SERVE_TRACE"] = trace
    try:
        eng.serve_start(max_walkers_in_flight=1 << 22, round_steps=round_steps, idle_poll_us=idle_poll_us)
        try:
            lat, steps, span, err = serve_stream(eng, jobs, gaps)
        finally:
            eng.serve_stop()
        out = {"queries": len(jobs)}
        if err:
            out.update({"completed": len(lat), "error": err})
        else:
            out.update({"p50_ms": round(float(np.percentile(lat, 50)), 3),
                        "p99_ms": round(float(np.percentile(lat, 99)), 3),
                        "steps_per_s": round(steps / span, 1), "occupancy": flight_occupancy(trace)})
            if groups:
                out["by_kind_ms"] = {k: {"p50": round(float(np.percentile(lat[m], 50)), 3),
                                         "p99": round(float(np.percentile(lat[m], 99)), 3)}
                                     for k, m in groups.items() if m.any()}
        return out
    finally:
        os.environ.pop("KKW_SERVE_TRACE", None)
        os.unlink(trace)


def serve_latency(eng, rp, rates, seconds, seed=7, idle_poll_us=0):
    """Poisson PPR queries (t=0.15, 1e3..1e5 walkers log-uniform, source =
    source of a uniformly drawn directed edge) at each rate; latency from
    submit to the query's last path on the host.  idle_poll_us: the serving
    loop's idle polling window (0: it sleeps until a query arrives)."""
    import knightking_amd as kk
    E = int(rp[-1])
    out = {}
    for rate in rates:
        rng = np.random.default_rng((seed, rate))   # one stream per rate, the same with and without polling
        n_q = int(rate * seconds)
        gaps = rng.exponential(1.0 / rate, n_q)
        srcs = np.searchsorted(rp, rng.integers(0, E, n_q), side="right") - 1
        walkers = np.round(10 ** (3 + 2 * rng.random(n_q))).astype(np.int64)
        jobs = [kk.ppr_job(int(walkers[i]), 0.15, True, seed=1000 + i,
                           start_vertices=np.array([srcs[i]], np.uint32)) for i in range(n_q)]
        out[str(rate)] = serve_session(eng, jobs, gaps, idle_poll_us=idle_poll_us)
        log(f"serving {rate} qps{' (idle polling %d us)' % idle_poll_us if idle_poll_us else ''}: {out[str(rate)]}")
        if "error" in out[str(rate)]:
            break
    return out


def flight_occupancy(trace_file):
    """From the serving loop's flight trace: time-weighted mean number of
    flights on the device (admission to completion seen) and the fraction of
    the busy span with all four streams occupied."""
    rows = np.loadtxt(trace_file, delimiter=",", skiprows=1, ndmin=2)
    if rows.size == 0:
        return None
    t0, t1 = rows[:, 1], rows[:, 3]   # admit_ms, phase1_seen_ms
    ev = np.concatenate([np.stack([t0, np.ones_like(t0)], 1), np.stack([t1, -np.ones_like(t1)], 1)])
    ev = ev[np.lexsort((ev[:, 1], ev[:, 0]))]
    level, last, area, full, span0 = 0, ev[0, 0], 0.0, 0.0, ev[0, 0]
    for t, d in ev:
        dt = t - last
        area += level * dt
        if level >= 4:
            full += dt
        level += int(d)
        last = t
    span = max(last - span0, 1e-9)
    return {"flights": int(len(rows)), "mean_flights_in_flight": round(float(area / span), 3),
            "frac_time_four_busy": round(float(full / span), 4),
            "mean_queries_per_flight": round(float(rows[:, 6].mean()), 2)}


def serve_mixed(eng, rp, rate, seconds, seed=8, round_steps=0):
    """Mixed queries at `rate`: kind uniform over PPR (t = 0.15, degree-
    proportional source), DeepWalk (L = 80) and node2vec (p = 2, q = 0.5,
    L = 80), all weighted, 1e3..1e5 walkers log-uniform; p50/p99 latency and
    how many of the loop's four flights were busy (from its flight trace)."""
    import knightking_amd as kk
    rng = np.random.default_rng(seed)
    E = int(rp[-1])
    n_q = int(rate * seconds)
    gaps = rng.exponential(1.0 / rate, n_q)
    walkers = np.round(10 ** (3 + 2 * rng.random(n_q))).astype(np.int64)
    kinds = rng.integers(0, 3, n_q)
    srcs = np.searchsorted(rp, rng.integers(0, E, n_q), side="right") - 1
    V = len(rp) - 1
    jobs = []
    for i in range(n_q):
        begin = int(rng.integers(0, V))
        if kinds[i] == 0:
            jobs.append(kk.ppr_job(int(walkers[i]), 0.15, True, seed=3000 + i,
                                   start_vertices=np.array([srcs[i]], np.uint32)))
        elif kinds[i] == 1:
            jobs.append(kk.deepwalk_job(int(walkers[i]), 80, True, seed=3000 + i, walker_begin=begin))
        else:
            jobs.append(kk.node2vec_job(int(walkers[i]), 80, 2.0, 0.5, True, seed=3000 + i, walker_begin=begin))
    out = {"rate_qps": rate, "kinds": "PPR t=0.15 / DeepWalk L=80 / node2vec L=80, weighted",
           "walkers": "10^(3+2U)", "round_steps": round_steps}
    groups = {"ppr": kinds == 0, "deepwalk": kinds == 1, "node2vec": kinds == 2}
    out.update(serve_session(eng, jobs, gaps, round_steps, groups))
    log(f"mixed serving {rate} qps: {out}")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--unweighted", action="store_true", help="node2vec without edge weights")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="use --shape instead of the autotuner's pick")
    ap.add_argument("--shape", default="64,16,64",
                    help="block,blocks_per_cu,chunk of the persistent launch with --no-autotune")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched-mixed and DeepWalk lines")
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--calib", action="store_true",
                    help="profiling pass: run the calibration copy only twice")
    ap.add_argument("--rebuilds", type=int, default=2,
                    help="rebuild the alias tables this many times and time one persistent launch on each "
                         "placement (the spread between table placements)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "the multi-GPU design is documented but not built", file=sys.stderr)
        sys.exit(2)

    import knightking_amd as kk

    weighted = not args.unweighted
    eng = kk.Engine()
    t0 = time.perf_counter()
    eng.graph_rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=1, weighted=True)
    graph_s = time.perf_counter() - t0
    V, E = eng.info()
    log(f"R-MAT-{args.scale}: V={V} E={E} built in {graph_s:.2f}s")
    eng.prepare()   # alias tables (weighted) and neighbour hash sets
    tp = eng.timing()
    alias_s = tp["alias_ms"] / 1e3 if weighted else 0.0
    index_s = tp["index_ms"] / 1e3
    prepare_s = tp["prepare_ms"] / 1e3   # both tables, built concurrently
    log(f"alias tables in {alias_s:.3f}s, neighbour hash sets in {index_s:.3f}s, together {prepare_s:.3f}s")

    job_args = dict(seed=1)
    job = kk.node2vec_job(V, 80, 2.0, 0.5, weighted, seed=1)

    # instrumented run: exact trial / search / line counts for the byte model
    sb = eng.batch([job], kk.MODE_PERSISTENT, kk.BATCH_STATS)
    sb.launch()
    sb.sync()
    t = eng.timing()
    sb.close()
    ref_bytes = reference_model_bytes(t, weighted)
    touched = kernel_bytes(t, V, LINE)
    bytes_launch = kernel_bytes(t, V, L2_LINE)   # algorithmic bytes: one L2 line per random access
    log(f"stats: steps={t['steps']} trials={t['trials']} searches={t['searches']} lines={t['search_lines']} "
        f"rows={t['row_loads']} ref_queries={t['ref_queries']} ref_lines={t['ref_query_lines']} "
        f"reference model bytes/step={ref_bytes / max(t['steps'], 1):.1f} "
        f"kernel bytes/step={bytes_launch / max(t['steps'], 1):.1f} (64 B lines: {touched / max(t['steps'], 1):.1f})")

    pb = eng.batch([job], kk.MODE_PERSISTENT, kk.BATCH_GRAPH)
    shape = tuple(int(x) for x in args.shape.split(","))
    if args.no_autotune:
        pb.set_shape(*shape)
    else:
        shape = pb.autotune(None)
        log(f"autotuned shape block/blocks_per_cu/chunk = {shape}")
    wall_p, kms_p = time_batch(pb, args.steps, args.warmup)
    steps_per_launch = pb.steps(0)
    assert steps_per_launch == t["steps"], "instrumented and timed runs disagree"
    value = steps_per_launch * args.steps / wall_p
    kernel_s = float(np.mean(kms_p)) / 1e3
    log(f"persistent: {wall_p / args.steps * 1e3:.2f} ms/launch wall, kernel {kernel_s * 1e3:.2f} ms, "
        f"{value / 1e9:.3f} G steps/s")

    plain = None
    if not args.no_plain:
        lb = eng.batch([job], kk.MODE_PLAIN, 0)
        wall_l, kms_l = time_batch(lb, args.steps, args.warmup)
        assert lb.steps(0) == steps_per_launch
        plain = steps_per_launch * args.steps / wall_l
        log(f"plain: {wall_l / args.steps * 1e3:.2f} ms/launch, {plain / 1e9:.3f} G steps/s")
        lb.close()

    copy_ms = eng.stream_copy_ms(CALIB_BYTES, 2 if args.calib else 10)
    copy_gbs = 2 * CALIB_BYTES / (copy_ms / 1e3) / 1e9
    log(f"streaming copy of {CALIB_BYTES >> 20} MiB: {copy_ms:.3f} ms, {copy_gbs:.0f} GB/s read+write")

    # ceiling for independent random 16-byte reads (one 128-byte line each),
    # against which the kernel's own rate of data-dependent random reads is set
    gather_lps = gather_small = None
    if not args.calib:
        gather_lps = eng.random_gather_lines_per_s(GATHER_BYTES)
        log(f"random 16-byte reads over {GATHER_BYTES >> 30} GiB: {gather_lps / 1e9:.2f} G lines/s "
            f"({gather_lps * L2_LINE / 1e9:.0f} GB/s of 128-byte lines)")
        # the same reads over 1 GiB, whose translations all hit the per-CU
        # translation caches: what the rate is without translation misses
        gather_small = eng.random_gather_lines_per_s(1 << 30)
        log(f"random 16-byte reads over 1 GiB: {gather_small / 1e9:.2f} G lines/s")
    accesses = t["row_loads"] + t["trials"] + t["search_lines"]
    access_rate = accesses / kernel_s

    # algorithmic bytes: one 64-byte transaction per data-dependent random
    # access + path and length bytes; the 128-byte L2-line figure beside it
    achieved = touched / kernel_s / 1e9
    achieved_128 = bytes_launch / kernel_s / 1e9
    # HBM traffic per launch: measured by rocprofv3 (FETCH_SIZE, WRITE_SIZE)
    # on a run of this workload, reported only when that profile was taken on
    # the same kernel and table-construction sources as this run
    traffic = None
    traffic_source = None
    tf = os.path.join(ROOT, "profiles", "node2vec_rmat24_hbm_traffic.json")
    if os.path.exists(tf) and weighted and args.scale == 24:
        with open(tf) as f:
            prof = json.load(f)
        if prof.get("source_digest") == source_digest():
            traffic = prof.get("hbm_bytes_per_launch")
            traffic_source = f"{prof.get('source')}: {prof.get('command')} (sources {prof['source_digest']})"
        else:
            traffic_source = (f"none: {tf} was profiled on sources {prof.get('source_digest')}, "
                              f"this run's are {source_digest()}")

    # the same launch on rebuilt tables: the library's device pool gives a
    # rebuilt table the block its predecessor's predecessor held, so these
    # launches see at most two table placements of this process
    placements = []
    for r in range(args.rebuilds):
        pb.close()
        eng.prepare()
        pb = eng.batch([job], kk.MODE_PERSISTENT, kk.BATCH_GRAPH)
        pb.set_shape(*shape)
        _, kms_r = time_batch(pb, 1, 1)
        placements.append(round(kms_r[0], 3))
    if placements:
        log(f"persistent launch on rebuilt tables: {placements} ms (first placement {kernel_s * 1e3:.3f} ms)")

    cpu = None
    parity = None
    if not args.no_cpu and weighted:
        rate, threads, sample, parity = cpu_baseline(eng, job_args, pb, V)
        hc = host_cpus()
        cpu = {"value": round(rate, 1), "unit": "walker steps/s", "cores": threads, "kind": "port",
               "sample": sample, "cpu_model": hc["cpu_model"], "nproc": hc["nproc"],
               "affinity_cpus": hc["affinity_cpus"], "cgroup_cpus": hc["cgroup_cpus"],
               "omp_num_threads": hc["omp_num_threads"],
               "threads_rule": "all CPUs the process may use (affinity, cgroup quota), capped by OMP_NUM_THREADS"}
        log(f"cpu baseline {rate / 1e6:.2f} M steps/s on {threads} threads; sample parity {parity}")
    run_extra = not args.no_extra and weighted and args.scale == 24
    builtin_paths = pb.paths(0) if run_extra else None
    pb.close()

    variants = None
    if not args.no_plain:
        # the same two launch forms with the binary search over sorted lists in
        # place of the neighbour hash sets
        eng.adj_index(False)
        variants = {}
        for name, mode in (("persistent_binary_search", kk.MODE_PERSISTENT), ("plain_binary_search", kk.MODE_PLAIN)):
            vb = eng.batch([job], mode, 0)
            wall_v, _ = time_batch(vb, args.steps, args.warmup)
            assert vb.steps(0) == steps_per_launch
            variants[name] = round(steps_per_launch * args.steps / wall_v, 1)
            vb.close()
            log(f"{name}: {variants[name] / 1e9:.3f} G steps/s")
        eng.adj_index(True)

    extra_batched = extra_deepwalk = extra_unweighted = extra_rmat14 = extra_program = None
    if run_extra:
        extra_program = program_line(kk, eng, V, builtin_paths)
        builtin_paths = None
        extra_unweighted = unweighted_line(kk, eng, V, args.steps, args.warmup)
        eng.close()
        extra_batched = batched_vs_plain(kk, args.steps, args.warmup)
        extra_deepwalk = deepwalk_line(kk, args.steps, args.warmup)
        extra_rmat14 = rmat14_line(kk, args.steps, args.warmup)
        eng = kk.Engine()
        eng.graph_rmat(args.scale, 16, 0.57, 0.19, 0.19, seed=1, weighted=True)
        eng.prepare()

    if cpu is not None and extra_rmat14 is not None:
        # the reference's CPU-runnable case timed whole on the same cores
        cpu["config1_whole_job"] = {
            "workload": "node2vec p=2 q=0.5 weighted, R-MAT scale 14, 16384 walkers (one per vertex), L=80, seed 1",
            "steps_per_s": extra_rmat14["cpu_oracle_steps_per_s"], "threads": extra_rmat14["cpu_threads"],
            "gpu_plain_steps_per_s": extra_rmat14["plain_steps_per_s"], "bit_exact": extra_rmat14["bit_exact"]}

    serving = None
    if not args.no_serve:
        rp, _, _, _ = eng.download(weights=False)
        serving = serve_latency(eng, rp, [200, 500, 1000, 2000], args.serve_seconds)
        # the same streams with the idle loop polling for 10 ms after its last
        # flight (it then admits an arrival without a wake-up)
        serving["idle_poll_10ms"] = serve_latency(eng, rp, [200, 2000], args.serve_seconds, idle_poll_us=10000)
        serving["mixed"] = [serve_mixed(eng, rp, r, args.serve_seconds, round_steps=k)
                            for r, k in ((2000, 0), (2000, 8), (8000, 0), (8000, 8))]

    line = {
        "metric": "walker steps/s, node2vec RMAT-24 1 GPU (batched vs plain); p99 PPR query latency",
        "value": round(value, 1),
        "unit": "walker steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_p / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32+f32",
        "data": "synthetic",
        "config": {
            "workload": f"node2vec p=2 q=0.5 {'weighted' if weighted else 'unweighted'}, undirected R-MAT "
                        f"scale {args.scale} (a/b/c/d=0.57/0.19/0.19/0.05, edge factor 16, dedup+mirror), "
                        f"walkers=|V|={V}, L=80, persistent work-queue kernel",
            "vertices": V, "edges": E, "walk_length": 80, "walkers": V,
            "steps_per_launch": int(steps_per_launch),
            "launch_shape": {"block": shape[0], "blocks_per_cu": shape[1], "chunk": shape[2]},
            "parallelism": "single GPU",
        },
        "plain_steps_per_s": None if plain is None else round(plain, 1),
        "persistent_over_plain": None if plain is None else round(value / plain, 3),
        "variants_steps_per_s": variants,
        "graph_build_s": round(graph_s, 3),
        "alias_build_s": round(alias_s, 3),
        "neighbour_set_build_s": round(index_s, 3),
        # the paper's end-to-end time: static tables built, then one walk of all walkers
        "prepare_s": round(prepare_s, 4),
        "end_to_end_s": round(prepare_s + kernel_s, 4),
        "end_to_end_steps_per_s": round(steps_per_launch / (prepare_s + kernel_s), 1),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_source,
            "algorithmic_bytes_per_launch": int(touched),
            "algorithmic_bytes_per_step": round(touched / max(steps_per_launch, 1), 1),
            "bytes_model": "one 64 B transaction per data-dependent random access (row pair at a walker start, "
                           "alias bucket, each distinct adjacency-probe block) + path and length bytes",
            "random_line_ceiling_frac": None if gather_lps is None else
            round(gather_lps * LINE / 1e9 / HBM_PEAK_GBS, 4),
            "lines_128B": {"bytes_per_launch": int(bytes_launch),
                           "bytes_per_step": round(bytes_launch / max(steps_per_launch, 1), 1),
                           "achieved": round(achieved_128, 1), "frac": round(achieved_128 / HBM_PEAK_GBS, 4)},
            "reference_model_bytes_per_step": round(ref_bytes / max(steps_per_launch, 1), 1),
            "reference_model_gbs": round(ref_bytes / kernel_s / 1e9, 1),
            "traffic_gbs": None if traffic is None else round(traffic / kernel_s / 1e9, 1),
            "kernel_ms": round(kernel_s * 1e3, 3),
            "trials_per_step": round(t["trials"] / max(t["steps"], 1), 4),
            "searches_per_step": round(t["searches"] / max(t["steps"], 1), 4),
            "search_lines_per_search": round(t["search_lines"] / max(t["searches"], 1), 3),
            "stream_copy_gbs": round(copy_gbs, 1),
            "random_accesses_per_s": round(access_rate, 1),
            "random_gather_lines_per_s": None if gather_lps is None else round(gather_lps, 1),
            "random_gather_1GiB_lines_per_s": None if gather_small is None else round(gather_small, 1),
            "placement_kernel_ms": [round(kernel_s * 1e3, 3)] + placements,
            "random_gather_gbs": None if gather_lps is None else round(gather_lps * L2_LINE / 1e9, 1),
            "frac_of_random_gather": None if gather_lps is None else round(access_rate / gather_lps, 4),
        },
        "cpu_baseline": cpu,
        "sample_parity": parity,
        "batched_mixed_jobs": extra_batched,
        "deepwalk_rmat22": extra_deepwalk,
        "node2vec_unweighted_rmat24": extra_unweighted,
        "node2vec_rmat14_cpu_case": extra_rmat14,
        "node2vec_program_rmat24": extra_program,
        "ppr_serving": serving,
    }
    print(json.dumps(line), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
