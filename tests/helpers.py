"""Shared checks of the GPU tests: a job descriptor run on the oracle, row
comparisons that stream through large results in chunks, edge existence."""
import numpy as np

import knightking_amd as kk
from oracle import oracle as O


def oracle_walk(g, job, walker_num=None, walker_begin=None, walker_stride=1, threads=0):
    """The oracle's paths for the kk.Job `job` (optionally a strided sample of
    its walkers: row i is walker walker_begin + i * walker_stride)."""
    return O.walk(g, job.kind, job.walker_num if walker_num is None else walker_num, job.walk_length,
                  job.use_weights, walker_begin=job.walker_begin if walker_begin is None else walker_begin,
                  terminate_prob=job.terminate_prob, p=job.p, q=job.q, seed=job.seed, start_mode=job.start_mode,
                  starts=job.starts, schemes=job.schemes, threads=threads, walker_stride=walker_stride)


def rows_equal(verts, lens, o_paths, o_lens, chunk=1 << 20):
    """verts[i, :lens[i]] == o_paths[i, :o_lens[i]] for every row, lengths equal;
    compared in chunks of rows (full-size results hold billions of entries)."""
    if not np.array_equal(np.asarray(lens), np.asarray(o_lens)):
        return False
    n = len(o_lens)
    w = min(verts.shape[1], o_paths.shape[1])
    if n and int(np.max(o_lens)) > w:
        return False
    col = np.arange(w)[None, :]
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        m = col < o_lens[a:b, None]
        if not np.all((verts[a:b, :w] == o_paths[a:b, :w]) | ~m):
            return False
    return True


def assert_rows_match_oracle(gpu, o_paths, o_lens, o_trunc, rows=None, what=""):
    """GPU result rows (all, or the listed rows) against oracle rows."""
    if rows is None:
        verts, lens, trunc = gpu.vertices, gpu.lengths, gpu.truncated
    else:
        verts, lens, trunc = gpu.vertices[rows], gpu.lengths[rows], gpu.truncated[rows]
    assert np.array_equal(lens, o_lens), f"{what}: lengths differ"
    assert np.array_equal(trunc, o_trunc), f"{what}: truncation flags differ"
    assert rows_equal(verts, lens, o_paths, o_lens), f"{what}: path entries differ"


def packed_matches_oracle(got, o_paths, o_lens, o_trunc):
    """A served result in packed form against oracle rows."""
    if not (np.array_equal(got.lengths, o_lens) and np.array_equal(got.truncated, o_trunc)):
        return False
    mask = np.arange(o_paths.shape[1])[None, :] < o_lens[:, None]
    return bool(np.array_equal(got.vertices, o_paths[mask]) and got.offsets[-1] == len(got.vertices))


def edge_keys(rp, col):
    """Sorted (u << 32 | v) keys of a CSR whose lists are sorted by neighbour."""
    deg = np.diff(rp.astype(np.int64))
    return (np.repeat(np.arange(len(rp) - 1, dtype=np.int64), deg) << 32) | col.astype(np.int64)


def edges_exist(key, u, v):
    probe = (np.asarray(u, np.int64) << 32) | np.asarray(v, np.int64)
    pos = np.searchsorted(key, probe)
    return bool(np.all(key[np.minimum(pos, len(key) - 1)] == probe))


def job_desc(job):
    return (f"kind {job.kind} walkers {job.walker_num} begin {job.walker_begin} L {job.walk_length} "
            f"weights {job.use_weights} seed {job.seed}")


__all__ = ["oracle_walk", "rows_equal", "assert_rows_match_oracle", "packed_matches_oracle", "edge_keys",
           "edges_exist", "job_desc", "kk", "O"]


class HipRuntime:
    """The HIP runtime libkkw itself is bound to (found in /proc/self/maps),
    through ctypes: device memory, pinned host memory and copies for tests
    that must not bring a second runtime (torch's) into the process."""

    def __init__(self):
        import ctypes as C
        with open("/proc/self/maps") as f:
            paths = {line.split()[-1] for line in f if "libamdhip64.so" in line}
        assert len(paths) == 1, f"expected the one HIP runtime libkkw uses, found {sorted(paths)}"
        lib = C.CDLL(paths.pop())
        P, Z, I = C.c_void_p, C.c_size_t, C.c_int
        for name, args in (("hipMalloc", [C.POINTER(P), Z]), ("hipFree", [P]), ("hipHostMalloc", [C.POINTER(P), Z, C.c_uint]),
                           ("hipHostFree", [P]), ("hipMemcpy", [P, P, Z, I]), ("hipMemcpyAsync", [P, P, Z, I, P]),
                           ("hipStreamSynchronize", [P]), ("hipDeviceSynchronize", [])):
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = I
        self.C = C
        self.lib = lib

    def host_alloc(self, nbytes):
        p = self.C.c_void_p()
        assert self.lib.hipHostMalloc(self.C.byref(p), max(nbytes, 16), 0) == 0
        return p.value

    def copy_async_d2h(self, dst, src, nbytes, stream):
        assert self.lib.hipMemcpyAsync(dst, src, nbytes, 2, stream) == 0
