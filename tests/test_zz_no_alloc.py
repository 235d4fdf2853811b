"""kkw_run_noalloc (walks into caller-owned device buffers, nothing allocated
on the device) against kkw_run on the same jobs: same lengths and rows, on two
calls in a row, the second with a smaller input, in both launch forms.

Device buffers come from the HIP runtime libkkw itself is bound to (found in
/proc/self/maps), through ctypes; results are read back with hipMemcpy."""
import ctypes as C
import math

import numpy as np
import pytest

import knightking_amd as kk

pytestmark = pytest.mark.gpu

HIP_D2H = 2


class HipMem:
    def __init__(self):
        with open("/proc/self/maps") as f:
            paths = {line.split()[-1] for line in f if "libamdhip64.so" in line}
        assert len(paths) == 1, f"expected the one HIP runtime libkkw uses, found {sorted(paths)}"
        lib = C.CDLL(paths.pop())
        lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        lib.hipMalloc.restype = C.c_int
        lib.hipFree.argtypes = [C.c_void_p]
        lib.hipFree.restype = C.c_int
        lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        lib.hipMemcpy.restype = C.c_int
        lib.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        lib.hipMemset.restype = C.c_int
        lib.hipDeviceSynchronize.restype = C.c_int
        self.lib = lib
        self.held = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(nbytes, 16)) == 0
        self.held.append(p.value)
        assert self.lib.hipMemset(p.value, 0xFF, max(nbytes, 16)) == 0
        return p.value

    def fill(self, ptr, nbytes):
        assert self.lib.hipMemset(ptr, 0xFF, max(nbytes, 16)) == 0
        assert self.lib.hipDeviceSynchronize() == 0

    def read(self, ptr, count, dtype=np.uint32):
        out = np.empty(count, dtype)
        if count:
            assert self.lib.hipMemcpy(out.ctypes.data, ptr, out.nbytes, HIP_D2H) == 0
        return out

    def free(self):
        for p in self.held:
            self.lib.hipFree(p)
        self.held = []


def capacity(job):
    """Path capacity in vertices: L + 1, or for PPR without a cap the library's
    default cap ceil(27.7 / t) + 1."""
    if job.kind == kk.PPR and job.walk_length == 0:
        return int(math.ceil(27.7 / float(np.float32(job.terminate_prob)))) + 1
    return job.walk_length + 1


def stride_for(job, mode):
    c = capacity(job)
    if mode == kk.MODE_PERSISTENT and c >= 32:
        return (c + 31) // 32 * 32   # whole 128-byte lines: the line-buffered path
    return (c + 3) // 4 * 4


@pytest.fixture(scope="module")
def eng():
    print("[gpu] no-alloc: R-MAT scale 12 weighted, 5 vertex types", flush=True)
    e = kk.Engine()
    e.graph_rmat(12, 16, seed=3, weighted=True)
    e.type_by_dst(5, seed=11)
    e.prepare()   # alias tables and hash sets; no uniform buckets (built only by kkw_run)
    mem = HipMem()
    yield e, mem
    mem.free()
    e.close()


def jobs_big():
    rng = np.random.default_rng(5)
    return [
        kk.node2vec_job(3000, 40, 2.0, 0.5, weighted=True, seed=21),
        kk.deepwalk_job(2000, 30, weighted=False, seed=22, walker_begin=100,
                        start_mode=kk.START_LIST, starts=rng.integers(0, 4096, 700).astype(np.uint32)),
        kk.ppr_job(2500, 0.1, weighted=True, seed=23, start_vertices=rng.integers(0, 4096, 50).astype(np.uint32)),
        kk.metapath_job(1800, 12, [[1, 2, 4], [3, 12], [31, 16, 8, 2]], weighted=True, seed=24),
        kk.node2vec_job(900, 9, 0.5, 2.0, weighted=False, seed=25, start_mode=kk.START_UNIFORM),
    ]


def jobs_small():
    return [
        kk.node2vec_job(500, 25, 2.0, 0.5, weighted=True, seed=31, walker_begin=7),
        kk.ppr_job(300, 0.15, weighted=False, seed=32, cap=40),
    ]


def run_both(e, mem, jobs, mode, bufs):
    """kkw_run_noalloc into bufs, then kkw_run of the same jobs; compare."""
    strides = [stride_for(j, mode) for j in jobs]
    for j, s, (pp, lp) in zip(jobs, strides, bufs):
        assert s * j.walker_num * 4 <= bufs_cap(bufs, pp)
        mem.fill(pp, j.walker_num * s * 4)
        mem.fill(lp, j.walker_num * 4)
    print(f"[gpu] kkw_run_noalloc: {len(jobs)} jobs mode {mode}", flush=True)
    e.run_noalloc(jobs, mode, [b[0] for b in bufs[:len(jobs)]], strides, [b[1] for b in bufs[:len(jobs)]])
    ref = e.run(jobs, mode=mode)
    for i, (j, s, r) in enumerate(zip(jobs, strides, ref)):
        n = j.walker_num
        raw = mem.read(bufs[i][1], n)
        assert np.array_equal(raw & 0x7FFFFFFF, r.lengths), f"job {i}: lengths differ"
        assert np.array_equal((raw >> 31).astype(bool), r.truncated), f"job {i}: truncation flags differ"
        rows = mem.read(bufs[i][0], n * s).reshape(n, s)
        w = min(s, r.vertices.shape[1])
        assert int(r.lengths.max(initial=0)) <= w
        mask = np.arange(w)[None, :] < r.lengths[:, None]
        assert np.array_equal(np.where(mask, rows[:, :w], 0), np.where(mask, r.vertices[:, :w], 0)), \
            f"job {i}: rows differ"
        assert int(((raw & 0x7FFFFFFF).astype(np.int64) - 1).sum()) == r.steps


_caps = {}


def bufs_cap(bufs, pp):
    return _caps[pp]


def make_bufs(mem, jobs, mode):
    out = []
    for j in jobs:
        nb = j.walker_num * stride_for(j, mode) * 4
        pp = mem.alloc(nb)
        _caps[pp] = nb
        out.append((pp, mem.alloc(j.walker_num * 4)))
    return out


@pytest.mark.parametrize("mode", [kk.MODE_PERSISTENT, kk.MODE_PLAIN])
def test_noalloc_matches_kkw_run_twice(eng, mode):
    e, mem = eng
    big = jobs_big()
    bufs = make_bufs(mem, big, kk.MODE_PERSISTENT)   # persistent strides are the larger ones
    run_both(e, mem, big, mode, bufs)
    # second call, smaller input, into the same buffers
    run_both(e, mem, jobs_small(), mode, bufs)
    # one job alone (the persistent single-job kernels)
    run_both(e, mem, jobs_small()[:1], mode, bufs)


def test_noalloc_rejects_what_does_not_fit(eng):
    e, mem = eng
    j = kk.node2vec_job(64, 20, 2.0, 0.5, weighted=True, seed=1)
    bufs = make_bufs(mem, [j], kk.MODE_PLAIN)
    # a stride below the path capacity
    with pytest.raises(kk.KkwError) as ex:
        e.run_noalloc([j], kk.MODE_PLAIN, [bufs[0][0]], [20], [bufs[0][1]])
    assert ex.value.code == kk.KKW_E_INVALID
    # a start list beyond the fixed scratch array (4 MiB)
    big = kk.ppr_job(64, 0.2, weighted=True, seed=2, cap=20,
                     start_vertices=np.zeros((4 << 20) // 4 + 1, np.uint32))
    with pytest.raises(kk.KkwError) as ex:
        e.run_noalloc([big], kk.MODE_PERSISTENT, [bufs[0][0]], [24], [bufs[0][1]])
    assert ex.value.code == kk.KKW_E_NOMEM
    # and the engine still runs afterwards
    run_both(e, mem, [j], kk.MODE_PERSISTENT, bufs)
