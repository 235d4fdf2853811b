"""Python view of the oracle (liborc.so, the CPU restatement of KnightKing's
random-walk hot path).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, which compare the GPU results with it.  The product
(knightking_amd/libkkw.so) never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborc.so")

DEEPWALK, PPR, NODE2VEC, METAPATH = 0, 1, 2, 3
START_EQUAL, START_LIST, START_UNIFORM = 0, 1, 2

_lib = None


class OrcGraph(C.Structure):
    _fields_ = [
        ("V", C.c_uint32), ("E", C.c_uint64), ("row_ptr", C.c_void_p), ("col", C.c_void_p),
        ("w", C.c_void_p), ("etype", C.c_void_p), ("bp", C.c_void_p), ("bprim", C.c_void_p),
        ("balias", C.c_void_p), ("bmeta", C.c_void_p), ("vsum", C.c_void_p),
    ]


class OrcJob(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("use_alias", C.c_int32), ("walker_begin", C.c_uint32),
        ("walker_num", C.c_uint32), ("walk_length", C.c_uint32), ("path_stride", C.c_uint32),
        ("seed", C.c_uint64), ("ppr_ext", C.c_float), ("n2v_upper", C.c_float),
        ("n2v_lower", C.c_float), ("n2v_inv_p", C.c_float), ("n2v_inv_q", C.c_float),
        ("n2v_overflow", C.c_float), ("start_mode", C.c_int32), ("n_starts", C.c_uint32),
        ("starts", C.c_void_p), ("n_schemes", C.c_uint32), ("scheme_len", C.c_void_p),
        ("scheme_mask", C.c_void_p), ("vmask", C.c_void_p), ("walker_stride", C.c_uint32),
    ]


class OrcStats(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("trials", C.c_uint64), ("searches", C.c_uint64),
                ("search_probes", C.c_uint64), ("truncated", C.c_uint64)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_philox.argtypes = [P, P, P]
        L.orc_u01.restype = C.c_float
        L.orc_u01.argtypes = [C.c_uint32]
        L.orc_bounded.restype = C.c_uint32
        L.orc_bounded.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_rmat_raw.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, P, P]
        L.orc_edge_weight.restype = C.c_float
        L.orc_edge_weight.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.orc_build_undirected.restype = C.c_uint64
        L.orc_build_undirected.argtypes = [C.c_uint32, P, P, C.c_uint64, C.c_uint64, C.c_int, P, P, P]
        L.orc_csr_from_edges.argtypes = [C.c_uint32, P, P, P, C.c_uint64, C.c_int, C.c_int, P, P, P]
        L.orc_vertex_types.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, P]
        L.orc_alias_build.argtypes = [C.c_uint32, P, P, P, P, P, P, P, P, P]
        L.orc_alias_build_local.argtypes = [C.c_uint32, P, P, P, P, P, P]
        L.orc_node2vec_constants.argtypes = [C.c_float, C.c_float] + [C.POINTER(C.c_float)] * 5
        L.orc_walk.argtypes = [C.POINTER(OrcGraph), C.POINTER(OrcJob), P, P, C.c_int, C.POINTER(OrcStats)]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_philox(_p(c), _p(k), _p(o))
    return o


def u01(u: int) -> float:
    return lib().orc_u01(u)


def bounded(lo: int, hi: int, n: int) -> int:
    return lib().orc_bounded(lo, hi, n)


def rmat_thresholds(a=0.57, b=0.19, c=0.19):
    def thr(x):
        v = np.floor(x * 4294967296.0)
        return 0xFFFFFFFF if v >= 4294967295.0 else int(v)
    return thr(a), thr(a + b), thr(a + b + c)


def rmat_raw(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    n = edge_factor << scale
    src = np.zeros(n, np.uint32)
    dst = np.zeros(n, np.uint32)
    ta, tab, tabc = rmat_thresholds(a, b, c)
    lib().orc_rmat_raw(scale, n, ta, tab, tabc, seed, _p(src), _p(dst))
    return src, dst


def build_undirected(V, src, dst, seed=1, weighted=True):
    n = len(src)
    rp = np.zeros(V + 1, np.uint64)
    col = np.zeros(max(2 * n, 1), np.uint32)
    w = np.zeros(max(2 * n, 1), np.float32)
    E = lib().orc_build_undirected(V, _p(np.ascontiguousarray(src, np.uint32)),
                                   _p(np.ascontiguousarray(dst, np.uint32)), n, seed, int(weighted),
                                   _p(rp), _p(col), _p(w))
    return rp, col[:E].copy(), (w[:E].copy() if weighted else None)


def rmat_graph(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, weighted=True):
    src, dst = rmat_raw(scale, edge_factor, a, b, c, seed)
    return build_undirected(1 << scale, src, dst, seed, weighted)


def edge_weight(seed, u, v) -> float:
    return lib().orc_edge_weight(seed, u, v)


def csr_from_edges(V, src, dst, w=None, undirected=False, sort_adj=False):
    n = len(src)
    m = 2 * n if undirected else n
    rp = np.zeros(V + 1, np.uint64)
    col = np.zeros(max(m, 1), np.uint32)
    wo = np.zeros(max(m, 1), np.float32) if w is not None else None
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    w = None if w is None else np.ascontiguousarray(w, np.float32)
    lib().orc_csr_from_edges(V, _p(src), _p(dst), _p(w), n, int(undirected), int(sort_adj), _p(rp), _p(col), _p(wo))
    return rp, col[:m].copy(), (None if wo is None else wo[:m].copy())


def vertex_types(V, ntypes, seed):
    vt = np.zeros(V, np.uint8)
    lib().orc_vertex_types(V, ntypes, seed, _p(vt))
    return vt


def alias_build(row_ptr, col, w, etype=None):
    V = len(row_ptr) - 1
    E = len(col)
    bp = np.zeros(max(E, 1), np.float32)
    prim = np.zeros(max(E, 1), np.uint32)
    al = np.zeros(max(E, 1), np.uint32)
    meta = np.zeros(max(E, 1), np.uint32)
    vs = np.zeros(V, np.float32)
    et = None if etype is None else np.ascontiguousarray(etype, np.uint8)
    lib().orc_alias_build(V, _p(row_ptr), _p(col), _p(w), _p(et), _p(bp), _p(prim), _p(al), _p(meta), _p(vs))
    return bp[:E], prim[:E], al[:E], meta[:E], vs


def vmask_of(row_ptr, etype):
    """Per vertex, the OR of (1 << type) over its out-edges (metapath.hpp:102-121)."""
    V = len(row_ptr) - 1
    vm = np.zeros(V, np.uint8)
    deg = np.diff(row_ptr.astype(np.int64))
    nz = np.nonzero(deg > 0)[0]
    if len(nz):
        bits = np.left_shift(np.uint8(1), etype.astype(np.uint8))
        vm[nz] = np.bitwise_or.reduceat(bits, row_ptr[:-1].astype(np.int64)[nz])
    return vm


def node2vec_constants(p, q):
    out = [C.c_float() for _ in range(5)]
    lib().orc_node2vec_constants(p, q, *[C.byref(o) for o in out])
    return [o.value for o in out]


def ppr_default_cap(t: float) -> int:
    """Step cap used when a PPR job gives none: (1-t)^cap <= 1e-12."""
    if not t > 0:
        return 0
    return int(min(np.ceil(27.7 / np.float64(np.float32(t))), 1048576))


@dataclass
class Graph:
    row_ptr: np.ndarray
    col: np.ndarray
    w: Optional[np.ndarray] = None
    etype: Optional[np.ndarray] = None
    alias: Optional[tuple] = None   # (p, prim, alias, meta, vsum)
    vmask: Optional[np.ndarray] = None   # metapath type masks, computed on first use

    @property
    def V(self):
        return len(self.row_ptr) - 1

    def with_alias(self):
        if self.alias is None:
            self.alias = alias_build(self.row_ptr, self.col, self.w, self.etype)
        return self


def walk(g: Graph, kind, walker_num, walk_length=0, use_weights=True, walker_begin=0, terminate_prob=0.0,
         p=1.0, q=1.0, seed=1, start_mode=START_EQUAL, starts=None, schemes=None, threads=0,
         walker_stride=1):
    """Run one job on the oracle; returns (paths [n, stride], lengths, truncated, stats)."""
    keep = []
    j = OrcJob()
    j.kind = kind
    j.use_alias = int(bool(use_weights))
    j.walker_begin = walker_begin
    j.walker_num = walker_num
    j.walker_stride = walker_stride
    L = walk_length
    if kind == PPR:
        j.ppr_ext = np.float32(1.0) - np.float32(terminate_prob)
        if L == 0:
            L = ppr_default_cap(terminate_prob)
    j.walk_length = L
    j.path_stride = L + 1
    j.seed = seed
    if kind == NODE2VEC:
        up, lo, ip, iq, ov = node2vec_constants(p, q)
        j.n2v_upper, j.n2v_lower, j.n2v_inv_p, j.n2v_inv_q, j.n2v_overflow = up, lo, ip, iq, ov
    j.start_mode = start_mode
    if starts is not None:
        s = np.ascontiguousarray(starts, np.uint32)
        keep.append(s)
        j.n_starts = len(s)
        j.starts = s.ctypes.data
    if schemes is not None:
        lens = np.array([len(s) for s in schemes], np.uint32)
        masks = np.zeros((len(schemes), 8), np.uint8)
        for i, s in enumerate(schemes):
            masks[i, :len(s)] = s
        masks = np.ascontiguousarray(masks.reshape(-1))
        if g.vmask is None:
            g.vmask = vmask_of(g.row_ptr, g.etype)
        vm = g.vmask
        keep += [lens, masks, vm]
        j.n_schemes = len(schemes)
        j.scheme_len = lens.ctypes.data
        j.scheme_mask = masks.ctypes.data
        j.vmask = vm.ctypes.data
    og = OrcGraph()
    og.V = g.V
    og.E = len(g.col)
    og.row_ptr = g.row_ptr.ctypes.data
    og.col = g.col.ctypes.data
    og.w = None if g.w is None else g.w.ctypes.data
    og.etype = None if g.etype is None else g.etype.ctypes.data
    if use_weights:
        g.with_alias()
        bp, prim, al, meta, vs = g.alias
        og.bp, og.bprim, og.balias, og.bmeta, og.vsum = (bp.ctypes.data, prim.ctypes.data, al.ctypes.data,
                                                         meta.ctypes.data, vs.ctypes.data)
    paths = np.zeros((walker_num, L + 1), np.uint32)
    lens = np.zeros(max(walker_num, 1), np.uint32)
    st = OrcStats()
    lib().orc_walk(C.byref(og), C.byref(j), _p(paths), _p(lens), threads, C.byref(st))
    raw = lens[:walker_num]
    return paths, raw & 0x7FFFFFFF, (raw >> 31).astype(bool), st


# ---- walk programs -----------------------------------------------------------
_ROOT = os.path.dirname(_HERE)
PROGRAM_ORACLE_SRC = os.path.join(_HERE, "program_oracle.cpp")


def alias_build_local(row_ptr, w):
    """Alias tables whose primary/alias are local edge indices (walk programs)."""
    V = len(row_ptr) - 1
    E = int(row_ptr[-1])
    bp = np.zeros(max(E, 1), np.float32)
    prim = np.zeros(max(E, 1), np.uint32)
    al = np.zeros(max(E, 1), np.uint32)
    vs = np.zeros(V, np.float32)
    lib().orc_alias_build_local(V, _p(np.ascontiguousarray(row_ptr, np.uint64)),
                                _p(np.ascontiguousarray(w, np.float32)), _p(bp), _p(prim), _p(al), _p(vs))
    return bp, prim, al, vs


def program_oracle(source_path: str, build_dir: str):
    """Compile the oracle's walk-program loop (program_oracle.cpp) with one
    program's source and load it; returns the ctypes library."""
    import hashlib
    import subprocess
    with open(source_path, "rb") as f:
        tag = hashlib.sha1(f.read() + open(PROGRAM_ORACLE_SRC, "rb").read()).hexdigest()[:12]
    out = os.path.join(build_dir, f"orc_program_{tag}.so")
    if not os.path.exists(out):
        os.makedirs(build_dir, exist_ok=True)
        cmd = ["g++", "-O2", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
               "-I", os.path.join(_ROOT, "include"), "-I", _HERE,
               f'-DKKW_PROGRAM_SOURCE="{os.path.abspath(source_path)}"',
               PROGRAM_ORACLE_SRC, "-x", "c", os.path.join(_HERE, "kkw_oracle.c"), "-o", out + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"program oracle does not compile:\n{r.stderr}")
        os.replace(out + ".tmp", out)
    L = C.CDLL(out)
    P, U32, U64, I = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    L.orc_program_walk.restype = I
    L.orc_program_walk.argtypes = [U32, P, P, P, P, I, P, P, P, P, P, P, U64, U32, U32, U32, U32, I, U32, P, I, P, P]
    return L


def program_walk(plib, row_ptr, col, w=None, etype=None, arrays=(), params=(), use_weights=False, seed=1,
                 walker_num=0, max_length=0, walker_begin=0, start_vertices=None, uniform_start=False,
                 second_order=False):
    """Run a walk program on the oracle; returns (paths [n, stride], lengths,
    truncated) with stride = max_length + 1 rounded up to 4."""
    V = len(row_ptr) - 1
    row_ptr = np.ascontiguousarray(row_ptr, np.uint64)
    col = np.ascontiguousarray(col, np.uint32)
    w = None if w is None else np.ascontiguousarray(w, np.float32)
    etype = None if etype is None else np.ascontiguousarray(etype, np.uint8)
    keep = [np.ascontiguousarray(a) if a is not None else None for a in arrays]
    arr = (C.c_void_p * 8)(*([None if a is None else a.ctypes.data for a in keep] + [None] * (8 - len(keep))))
    par = (C.c_float * 8)(*([float(np.float32(x)) for x in params] + [0.0] * (8 - len(params))))
    bp = prim = al = vs = None
    if use_weights:
        bp, prim, al, vs = alias_build_local(row_ptr, w)
    stride = (max_length + 1 + 3) & ~3
    paths = np.zeros((max(walker_num, 1), stride), np.uint32)
    lens = np.zeros(max(walker_num, 1), np.uint32)
    if start_vertices is not None:
        starts = np.ascontiguousarray(start_vertices, np.uint32)
        mode = START_LIST
    else:
        starts = None
        mode = START_UNIFORM if uniform_start else START_EQUAL
    d = np.diff(col.astype(np.int64))
    starts_of = row_ptr[1:-1].astype(np.int64) - 1
    breaks = np.zeros(len(d), bool)
    breaks[starts_of[(starts_of >= 0) & (starts_of < len(d))]] = True
    sorted_lists = bool(np.all((d >= 0) | breaks))
    plib.orc_program_walk(V, _p(row_ptr), _p(col), _p(w), _p(etype), int(sorted_lists), arr, par,
                          _p(bp), _p(prim), _p(al), _p(vs),
                          seed, walker_begin, walker_num, max_length, stride, mode,
                          0 if starts is None else len(starts), _p(starts), int(second_order), _p(paths), _p(lens))
    raw = lens[:walker_num]
    return paths[:walker_num], raw & 0x7FFFFFFF, (raw >> 31).astype(bool)
