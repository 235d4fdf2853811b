"""knightking_amd -- MI355X random-walk engine (Python view of libkkw's C ABI).

The product is libkkw.so (HIP kernels + C ABI, include/kkw.h).  This module
loads it with ctypes and mirrors the reference's walk entry points
(KnightKing src/apps/node2vec.hpp:108, deepwalk.hpp:30, ppr.hpp:30,
metapath.hpp:70; WalkEngine::random_walk include/walk.hpp:423-436) for tests and
the benchmark.  There is no CPU fallback: without the built library, or
without a GPU, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libkkw.so")

KKW_OK = 0
KKW_E_INVALID, KKW_E_NOGRAPH, KKW_E_NOMEM, KKW_E_DEVICE = -1, -2, -3, -4
KKW_E_BUSY, KKW_E_CANCELLED, KKW_E_TIMEOUT, KKW_E_IO = -5, -6, -7, -8

DEEPWALK, PPR, NODE2VEC, METAPATH = 0, 1, 2, 3
START_EQUAL, START_LIST, START_UNIFORM = 0, 1, 2
MODE_PLAIN, MODE_PERSISTENT = 0, 1
OUT_HOST, OUT_DEVICE = 0, 1
FLAG_STATS = 1
BATCH_GRAPH, BATCH_STATS = 1, 2


class KkwError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"kkw error {code}: {msg}")
        self.code = code


class _Opts(C.Structure):
    _fields_ = [("device", C.c_int32), ("verbose", C.c_int32), ("hbm_budget", C.c_uint64)]


class _Job(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("use_weights", C.c_int32), ("walker_begin", C.c_uint32),
        ("walker_num", C.c_uint32), ("walk_length", C.c_uint32), ("terminate_prob", C.c_float),
        ("p", C.c_float), ("q", C.c_float), ("seed", C.c_uint64), ("start_mode", C.c_int32),
        ("n_starts", C.c_uint32), ("starts", C.POINTER(C.c_uint32)), ("n_schemes", C.c_uint32),
        ("scheme_len", C.POINTER(C.c_uint32)), ("scheme_mask", C.POINTER(C.c_uint8)),
    ]


class _Paths(C.Structure):
    _fields_ = [
        ("walker_begin", C.c_uint32), ("walker_num", C.c_uint32), ("stride", C.c_uint32),
        ("on_device", C.c_uint32), ("vertices", C.POINTER(C.c_uint32)),
        ("lengths", C.POINTER(C.c_uint32)), ("steps", C.c_uint64), ("truncated", C.c_uint64),
        ("offsets", C.POINTER(C.c_uint64)), ("priv", C.c_void_p),
    ]


class _ProgJob(C.Structure):
    _fields_ = [
        ("walker_begin", C.c_uint32), ("walker_num", C.c_uint32), ("max_length", C.c_uint32),
        ("use_weights", C.c_int32), ("seed", C.c_uint64), ("start_mode", C.c_int32), ("n_starts", C.c_uint32),
        ("starts", C.POINTER(C.c_uint32)), ("second_order", C.c_int32), ("mode", C.c_int32),
        ("param", C.c_float * 8),
    ]


class _Timing(C.Structure):
    _fields_ = [
        ("kernel_ms", C.c_double), ("total_ms", C.c_double), ("alias_ms", C.c_double),
        ("graph_ms", C.c_double), ("index_ms", C.c_double), ("steps", C.c_uint64), ("trials", C.c_uint64),
        ("searches", C.c_uint64), ("search_probes", C.c_uint64), ("row_loads", C.c_uint64),
        ("search_lines", C.c_uint64), ("ref_queries", C.c_uint64), ("ref_query_lines", C.c_uint64),
        ("prepare_ms", C.c_double), ("envelope_violations", C.c_uint64),
    ]


class _ServeOpts(C.Structure):
    _fields_ = [("max_walkers_in_flight", C.c_uint32), ("round_steps", C.c_uint32),
                ("queue_capacity", C.c_uint32), ("idle_poll_us", C.c_uint32)]


# kkw_epoch_consumer: (paths, stream, user) -> 0 to go on
EPOCH_CONSUMER = C.CFUNCTYPE(C.c_int, C.POINTER(_Paths), C.c_void_p, C.c_void_p)

# Every symbol include/kkw.h declares (tests check the library exports them all).
EXPORTED = [
    "kkw_last_error", "kkw_abi_version", "kkw_engine_create", "kkw_engine_destroy",
    "kkw_graph_from_edges", "kkw_graph_from_arrays", "kkw_graph_sort_adj", "kkw_graph_load_file", "kkw_graph_from_csr", "kkw_graph_rmat",
    "kkw_graph_set_edge_types", "kkw_graph_type_by_dst", "kkw_graph_info", "kkw_graph_download",
    "kkw_prepare", "kkw_graph_adj_index", "kkw_alias_download", "kkw_uniform_start", "kkw_run", "kkw_run_noalloc",
    "kkw_run_epochs", "kkw_paths_download", "kkw_paths_free",
    "kkw_paths_dump",
    "kkw_last_timing", "kkw_batch_create", "kkw_batch_set_shape", "kkw_batch_launch",
    "kkw_batch_sync", "kkw_batch_paths", "kkw_batch_destroy", "kkw_autotune", "kkw_diag_stream_copy",
    "kkw_diag_random_gather",
    "kkw_program_create", "kkw_program_destroy", "kkw_program_log", "kkw_program_set_array", "kkw_program_run",
    "kkw_serve_start", "kkw_submit", "kkw_poll", "kkw_wait", "kkw_cancel", "kkw_serve_stop",
]

_lib = None


def hip_runtimes(maps_text: Optional[str] = None) -> List[str]:
    """Distinct HIP runtime libraries (libamdhip64) mapped into this process,
    from /proc/self/maps (or the given text of such a file)."""
    if maps_text is None:
        try:
            with open("/proc/self/maps") as f:
                maps_text = f.read()
        except OSError:
            return []
    paths = set()
    for line in maps_text.splitlines():
        parts = line.split()
        if len(parts) >= 6 and "libamdhip64.so" in os.path.basename(parts[-1]):
            paths.add(parts[-1])
    return sorted(paths)


def check_one_hip_runtime(where: str, maps_text: Optional[str] = None) -> None:
    """Raise before any GPU call when two HIP runtimes are mapped: torch ships
    its own libamdhip64, and a process that loaded libkkw (with the system
    runtime) before torch initialised its GPU ends up with both, each driving
    the GPU with its own state; GPU calls through the second one have crashed
    (a segmentation fault inside a torch allocation).  Import torch and call
    torch.cuda.init() before the first Engine, so libkkw binds to torch's."""
    rt = hip_runtimes(maps_text)
    if len(rt) > 1:
        raise RuntimeError(f"{where}: two HIP runtimes are mapped in this process ({', '.join(rt)}); "
                           "import torch and call torch.cuda.init() before the first Engine, or keep torch "
                           "out of the process")


def load_library(path: str = LIB_PATH):
    """Load libkkw.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with __graft_entry__.build() "
                           "(the engine has no CPU fallback)")
    lib = C.CDLL(path)
    P, U32, U64, I32, D = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_double
    sig = {
        "kkw_last_error": (C.c_char_p, []),
        "kkw_abi_version": (I32, []),
        "kkw_engine_create": (I32, [C.POINTER(_Opts), C.POINTER(P)]),
        "kkw_engine_destroy": (I32, [P]),
        "kkw_graph_from_edges": (I32, [P, U32, P, U64, I32, I32, I32]),
        "kkw_graph_from_arrays": (I32, [P, U32, P, P, P, P, U64, I32, I32]),
        "kkw_graph_sort_adj": (I32, [P]),
        "kkw_graph_load_file": (I32, [P, U32, C.c_char_p, I32, I32, I32, I32]),
        "kkw_graph_from_csr": (I32, [P, U32, P, P, P]),
        "kkw_graph_rmat": (I32, [P, I32, I32, D, D, D, U64, I32]),
        "kkw_graph_set_edge_types": (I32, [P, P]),
        "kkw_graph_type_by_dst": (I32, [P, I32, U64]),
        "kkw_graph_info": (I32, [P, C.POINTER(U32), C.POINTER(U64)]),
        "kkw_graph_download": (I32, [P, P, P, P, P]),
        "kkw_prepare": (I32, [P]),
        "kkw_graph_adj_index": (I32, [P, I32]),
        "kkw_alias_download": (I32, [P, P, P, P, P, P]),
        "kkw_run": (I32, [P, C.POINTER(_Job), I32, I32, I32, I32, C.POINTER(_Paths)]),
        "kkw_run_noalloc": (I32, [P, C.POINTER(_Job), I32, I32, P, P, P]),
        "kkw_uniform_start": (U32, [U64, U32, U32]),
        "kkw_run_epochs": (I32, [P, C.POINTER(_Job), U32, I32, EPOCH_CONSUMER, P]),
        "kkw_paths_download": (I32, [C.POINTER(_Paths), C.POINTER(_Paths)]),
        "kkw_paths_free": (I32, [C.POINTER(_Paths)]),
        "kkw_paths_dump": (I32, [C.POINTER(_Paths), C.c_char_p, C.c_char_p, I32]),
        "kkw_last_timing": (I32, [P, C.POINTER(_Timing)]),
        "kkw_batch_create": (I32, [P, C.POINTER(_Job), I32, I32, I32, C.POINTER(P)]),
        "kkw_batch_set_shape": (I32, [P, I32, I32, I32]),
        "kkw_batch_launch": (I32, [P]),
        "kkw_batch_sync": (I32, [P, C.POINTER(D)]),
        "kkw_batch_paths": (I32, [P, I32, I32, C.POINTER(_Paths)]),
        "kkw_batch_destroy": (I32, [P]),
        "kkw_autotune": (I32, [P, C.c_char_p, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]),
        "kkw_diag_stream_copy": (I32, [P, U64, I32, C.POINTER(D)]),
        "kkw_diag_random_gather": (I32, [P, U64, I32, I32, C.POINTER(D)]),
        "kkw_program_create": (I32, [P, C.c_char_p, C.c_char_p, C.POINTER(P)]),
        "kkw_program_destroy": (I32, [P]),
        "kkw_program_log": (C.c_char_p, [P]),
        "kkw_program_set_array": (I32, [P, I32, P, U64]),
        "kkw_program_run": (I32, [P, C.POINTER(_ProgJob), C.POINTER(_Paths)]),
        "kkw_serve_start": (I32, [P, C.POINTER(_ServeOpts)]),
        "kkw_submit": (I32, [P, C.POINTER(_Job), C.POINTER(U64)]),
        "kkw_poll": (I32, [P, U64, C.POINTER(I32)]),
        "kkw_wait": (I32, [P, U64, D, C.POINTER(_Paths), C.POINTER(D)]),
        "kkw_cancel": (I32, [P, U64]),
        "kkw_serve_stop": (I32, [P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _check(rc: int):
    if rc != KKW_OK:
        msg = _lib.kkw_last_error().decode(errors="replace")
        raise KkwError(rc, msg)


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


@dataclass
class Job:
    """One walk job (kkw_job): the arguments of node2vec()/deepwalk()/ppr()/metapath()."""
    kind: int
    walker_num: int
    walk_length: int = 0
    use_weights: bool = False
    walker_begin: int = 0
    terminate_prob: float = 0.0
    p: float = 1.0
    q: float = 1.0
    seed: int = 1
    start_mode: int = START_EQUAL
    starts: Optional[np.ndarray] = None
    schemes: Optional[List[List[int]]] = None  # per scheme: allowed-type bitmask per state
    _keep: list = field(default_factory=list, repr=False)

    def to_c(self) -> _Job:
        j = _Job()
        j.kind = self.kind
        j.use_weights = 1 if self.use_weights else 0
        j.walker_begin = self.walker_begin
        j.walker_num = self.walker_num
        j.walk_length = self.walk_length
        j.terminate_prob = self.terminate_prob
        j.p, j.q = self.p, self.q
        j.seed = self.seed
        j.start_mode = self.start_mode
        self._keep = []
        if self.starts is not None:
            s = np.ascontiguousarray(self.starts, dtype=np.uint32)
            self._keep.append(s)
            j.n_starts = len(s)
            j.starts = s.ctypes.data_as(C.POINTER(C.c_uint32))
        if self.schemes is not None:
            lens = np.array([len(s) for s in self.schemes], dtype=np.uint32)
            masks = np.zeros((len(self.schemes), 8), dtype=np.uint8)
            for i, s in enumerate(self.schemes):
                masks[i, :len(s)] = s
            masks = np.ascontiguousarray(masks.reshape(-1))
            self._keep += [lens, masks]
            j.n_schemes = len(self.schemes)
            j.scheme_len = lens.ctypes.data_as(C.POINTER(C.c_uint32))
            j.scheme_mask = masks.ctypes.data_as(C.POINTER(C.c_uint8))
        return j


def uniform_start(seed: int, walker: int, v_num: int) -> int:
    """Start vertex of `walker` under START_UNIFORM for a job seeded `seed`
    (kkw_uniform_start: the kernels' draw, evaluated on the host)."""
    return int(load_library().kkw_uniform_start(seed, walker, v_num))


def node2vec_constants(p: float, q: float) -> List[float]:
    """node2vec's bounds and dynamic components as the engine computes them
    for a job: upper = max(1, 1/q) and lower = min(1/p, 1, 1/q) in double,
    stored as float (src/apps/node2vec.hpp:126-127); 1/p and 1/q in float
    (node2vec.hpp:182-191); the outlier overflow 1/p - upper in double, stored
    as float (node2vec.hpp:136).  Returns [upper, lower, 1/p, 1/q, overflow],
    the parameters 1-5 of the node2vec walk program (programs/node2vec.cpp)."""
    p32, q32 = np.float32(p), np.float32(q)
    ip, iq = 1.0 / float(p32), 1.0 / float(q32)
    upper = np.float32(max(1.0, iq))
    lower = np.float32(min(ip, min(1.0, iq)))
    inv_p = np.float32(1.0) / p32
    inv_q = np.float32(1.0) / q32
    overflow = np.float32(ip - float(upper))
    return [float(upper), float(lower), float(inv_p), float(inv_q), float(overflow)]


def node2vec_job(walker_num, walk_length, p, q, weighted=True, seed=1, **kw) -> Job:
    return Job(NODE2VEC, walker_num, walk_length, use_weights=weighted, p=p, q=q, seed=seed, **kw)


def deepwalk_job(walker_num, walk_length, weighted=True, seed=1, **kw) -> Job:
    return Job(DEEPWALK, walker_num, walk_length, use_weights=weighted, seed=seed, **kw)


def ppr_job(walker_num, terminate_prob, weighted=True, seed=1, start_vertices=None, cap=0, **kw) -> Job:
    mode = START_LIST if start_vertices is not None else START_EQUAL
    return Job(PPR, walker_num, cap, use_weights=weighted, terminate_prob=terminate_prob, seed=seed,
               start_mode=mode, starts=start_vertices, **kw)


def metapath_job(walker_num, walk_length, schemes, weighted=True, seed=1, **kw) -> Job:
    return Job(METAPATH, walker_num, walk_length, use_weights=weighted, schemes=schemes, seed=seed, **kw)


@dataclass
class Paths:
    """PathSet of one job, one segment, walker-id order (include/path.hpp:46-84)."""
    walker_begin: int
    vertices: np.ndarray      # [walker_num, stride] uint32
    lengths: np.ndarray       # [walker_num] uint32 (vertices per path)
    truncated: np.ndarray     # [walker_num] bool
    steps: int

    def path(self, i: int) -> np.ndarray:
        return self.vertices[i, : self.lengths[i]]


def _take_paths(p: _Paths) -> Paths:
    n, s = p.walker_num, p.stride
    raw = np.ctypeslib.as_array(p.lengths, shape=(n,)).copy() if n else np.zeros(0, np.uint32)
    lens = raw & 0x7FFFFFFF
    if n and bool(p.offsets):
        # packed form: row i at vertices + offsets[i] (offsets need not start
        # at 0); expand to rows of `stride` entries (unused entries zero)
        offs = np.ctypeslib.as_array(p.offsets, shape=(n + 1,)).astype(np.int64)
        lo, hi = int(offs[0]), int(offs[-1])
        flat = np.ctypeslib.as_array(p.vertices, shape=(max(hi, 1),))[lo:hi]
        verts = np.zeros((n, s), np.uint32)
        cols = np.arange(hi - lo) - np.repeat(offs[:-1] - lo, lens.astype(np.int64))
        verts[np.repeat(np.arange(n), lens.astype(np.int64)), cols] = flat
    elif n:
        verts = np.ctypeslib.as_array(p.vertices, shape=(n * s,)).reshape(n, s).copy()
    else:
        verts = np.zeros((0, s), np.uint32)
    out = Paths(p.walker_begin, verts, lens, (raw >> 31).astype(bool), int(p.steps))
    _lib.kkw_paths_free(C.byref(p))
    return out


@dataclass
class PackedPaths:
    """A PathSet in packed form: walker i's path is
    vertices[offsets[i]:offsets[i+1]] (lengths[i] entries)."""
    walker_begin: int
    vertices: np.ndarray      # [total] uint32
    offsets: np.ndarray       # [walker_num + 1] int64, offsets[0] == 0
    lengths: np.ndarray       # [walker_num] uint32
    truncated: np.ndarray     # [walker_num] bool
    steps: int

    def path(self, i: int) -> np.ndarray:
        return self.vertices[self.offsets[i]:self.offsets[i + 1]]


def _take_packed(p: _Paths) -> "PackedPaths":
    n = p.walker_num
    raw = np.ctypeslib.as_array(p.lengths, shape=(n,)).copy() if n else np.zeros(0, np.uint32)
    lens = raw & 0x7FFFFFFF
    if n and bool(p.offsets):
        offs = np.ctypeslib.as_array(p.offsets, shape=(n + 1,)).astype(np.int64)
        lo, hi = int(offs[0]), int(offs[-1])
        verts = np.ctypeslib.as_array(p.vertices, shape=(max(hi, 1),))[lo:hi].copy()
        offs = offs - lo
    elif n:   # row-major result: gather each row's entries
        s = p.stride
        rows = np.ctypeslib.as_array(p.vertices, shape=(n * s,)).reshape(n, s)
        offs = np.zeros(n + 1, np.int64)
        offs[1:] = np.cumsum(lens.astype(np.int64))
        verts = rows[np.arange(s)[None, :] < lens[:, None]].copy()
    else:
        offs = np.zeros(1, np.int64)
        verts = np.zeros(0, np.uint32)
    out = PackedPaths(p.walker_begin, verts, offs, lens, (raw >> 31).astype(bool), int(p.steps))
    _lib.kkw_paths_free(C.byref(p))
    return out


class Program:
    """A walk program (kkw_program): a first-order transition written as device
    functions (include/kkw_program.h), compiled at run time for the engine's GPU."""

    def __init__(self, eng: "Engine", source: str, options: Optional[str] = None):
        self.eng = eng
        h = C.c_void_p()
        _check(_lib.kkw_program_create(eng.h, source.encode(), None if options is None else options.encode(),
                                       C.byref(h)))
        self.h = h
        self._arrays = {}
        eng._children.add(self)

    @property
    def log(self) -> str:
        return _lib.kkw_program_log(self.h).decode(errors="replace")

    def set_array(self, slot: int, data: Optional[np.ndarray]):
        """kkw_ctx.array[slot] = a device copy of `data` (None clears it)."""
        if data is None:
            _check(_lib.kkw_program_set_array(self.h, slot, None, 0))
            self._arrays.pop(slot, None)
            return
        data = np.ascontiguousarray(data)
        _check(_lib.kkw_program_set_array(self.h, slot, _ptr(data), data.nbytes))
        self._arrays[slot] = data

    def run(self, walker_num: int, max_length: int, use_weights: bool = False, seed: int = 1,
            walker_begin: int = 0, start_vertices: Optional[np.ndarray] = None, uniform_start: bool = False,
            second_order: bool = False, params: Sequence[float] = (), mode: int = MODE_PLAIN) -> Paths:
        """Walk with the program; mode MODE_PLAIN (one lane per walker) or
        MODE_PERSISTENT (lanes refilled from a work queue): the same paths."""
        j = _ProgJob()
        j.mode = mode
        j.walker_begin, j.walker_num, j.max_length = walker_begin, walker_num, max_length
        j.use_weights, j.seed = int(use_weights), seed
        starts = None
        if start_vertices is not None:
            starts = np.ascontiguousarray(start_vertices, np.uint32)
            j.start_mode, j.n_starts = START_LIST, len(starts)
            j.starts = starts.ctypes.data_as(C.POINTER(C.c_uint32))
        else:
            j.start_mode = START_UNIFORM if uniform_start else START_EQUAL
        j.second_order = int(second_order)
        if len(params) > 8:
            raise ValueError("at most 8 parameters")
        for k, v in enumerate(params):
            j.param[k] = float(np.float32(v))
        out = _Paths()
        _check(_lib.kkw_program_run(self.h, C.byref(j), C.byref(out)))
        return _take_paths(out)

    def close(self):
        if getattr(self, "h", None):
            _lib.kkw_program_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _DeviceArray:
    """int32 device array in the __cuda_array_interface__ form torch.as_tensor
    reads; holds `owner` so the memory outlives the tensor."""

    def __init__(self, ptr: int, shape, owner):
        self._owner = owner
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": "<i4", "data": (ptr, False),
                                         "version": 2, "strides": None}


class Batch:
    """Resident batch (kkw_batch): job table, path buffers, optional hipGraph."""

    def __init__(self, eng: "Engine", jobs: Sequence[Job], mode: int, flags: int = 0):
        self.eng = eng
        self._cj = (_Job * len(jobs))(*[j.to_c() for j in jobs])
        self._jobs = list(jobs)
        h = C.c_void_p()
        _check(_lib.kkw_batch_create(eng.h, self._cj, len(jobs), mode, flags, C.byref(h)))
        self.h = h
        eng._children.add(self)

    def set_shape(self, block=256, blocks_per_cu=8, chunk=64):
        _check(_lib.kkw_batch_set_shape(self.h, block, blocks_per_cu, chunk))

    def launch(self):
        _check(_lib.kkw_batch_launch(self.h))

    def sync(self) -> float:
        ms = C.c_double()
        _check(_lib.kkw_batch_sync(self.h, C.byref(ms)))
        return ms.value

    def paths(self, i: int) -> Paths:
        p = _Paths()
        _check(_lib.kkw_batch_paths(self.h, i, OUT_HOST, C.byref(p)))
        return _take_paths(p)

    def steps(self, i: int) -> int:
        p = _Paths()
        _check(_lib.kkw_batch_paths(self.h, i, OUT_DEVICE, C.byref(p)))
        return int(p.steps)

    def paths_torch(self, i: int):
        """Job i's paths where the kernel wrote them, as torch tensors on the
        engine's GPU sharing the batch's memory (no copy): vertices int32
        [walker_num, stride] (ids read as int32), lengths int32 [walker_num]
        (bit 31 marks a capped path, so such lengths read negative).  For a
        consumer on the GPU (e.g. skip-gram pairs).  The tensors keep the
        batch object alive, hold the last launch's paths until the next
        launch, and must not be used once the batch or its engine is closed.

        torch bundles its own HIP runtime; one process can drive the GPU through
        only one of the two, so import torch and initialise its GPU before the
        first Engine: libkkw then binds to the runtime torch loaded."""
        import torch
        check_one_hip_runtime("Batch.paths_torch")
        if not torch.cuda.is_available():
            raise RuntimeError("torch does not see the GPU in this process: import torch and call "
                               "torch.cuda.init() before the first Engine, so that libkkw and torch "
                               "share one HIP runtime")
        p = _Paths()
        _check(_lib.kkw_batch_paths(self.h, i, OUT_DEVICE, C.byref(p)))
        n, s = int(p.walker_num), int(p.stride)
        dev = torch.device("cuda", self.eng.device)
        if n == 0:
            return (torch.empty((0, s), dtype=torch.int32, device=dev), torch.empty(0, dtype=torch.int32, device=dev))
        verts = _DeviceArray(C.cast(p.vertices, C.c_void_p).value, (n, s), self)
        lens = _DeviceArray(C.cast(p.lengths, C.c_void_p).value, (n,), self)
        return torch.as_tensor(verts, device=dev), torch.as_tensor(lens, device=dev)

    def autotune(self, cache_file: Optional[str] = None):
        b, bp, ch = C.c_int(), C.c_int(), C.c_int()
        _check(_lib.kkw_autotune(self.h, cache_file.encode() if cache_file else None,
                                 C.byref(b), C.byref(bp), C.byref(ch)))
        return b.value, bp.value, ch.value

    def close(self):
        if getattr(self, "h", None):
            _lib.kkw_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """WalkEngine on one GPU (include/walk.hpp:365-437)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, verbose: bool = False):
        load_library()
        check_one_hip_runtime("Engine()")
        o = _Opts(device, 1 if verbose else 0, hbm_budget)
        h = C.c_void_p()
        _check(_lib.kkw_engine_create(C.byref(o), C.byref(h)))
        self.h = h
        self.device = device
        # batches and programs hold the engine's handle: closed before it is
        self._children = weakref.WeakSet()

    def close(self):
        if getattr(self, "h", None):
            for child in list(getattr(self, "_children", ())):
                child.close()
            _lib.kkw_engine_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- graphs --------------------------------------------------------
    def graph_rmat(self, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, weighted=True):
        _check(_lib.kkw_graph_rmat(self.h, scale, edge_factor, a, b, c, seed, 1 if weighted else 0))

    def graph_from_edges(self, v_num, src, dst, w=None, make_undirected=False, sort_adj=False):
        n = len(src)
        if w is None:
            rec = np.zeros(n, dtype=[("s", "<u4"), ("d", "<u4")])
        else:
            rec = np.zeros(n, dtype=[("s", "<u4"), ("d", "<u4"), ("w", "<f4")])
            rec["w"] = w
        rec["s"], rec["d"] = src, dst
        _check(_lib.kkw_graph_from_edges(self.h, v_num, _ptr(rec), n, 0 if w is None else 1,
                                         1 if make_undirected else 0, 1 if sort_adj else 0))

    def graph_from_arrays(self, v_num, src, dst, w=None, etype=None, make_undirected=False, sort_adj=False):
        src = np.ascontiguousarray(src, np.uint32)
        dst = np.ascontiguousarray(dst, np.uint32)
        w = None if w is None else np.ascontiguousarray(w, np.float32)
        etype = None if etype is None else np.ascontiguousarray(etype, np.uint8)
        _check(_lib.kkw_graph_from_arrays(self.h, v_num, _ptr(src), _ptr(dst), _ptr(w), _ptr(etype), len(src),
                                          int(make_undirected), int(sort_adj)))

    def sort_adj(self):
        _check(_lib.kkw_graph_sort_adj(self.h))

    def load_graph(self, v_num, path, text=False, weighted=False, make_undirected=False, sort_adj=False):
        _check(_lib.kkw_graph_load_file(self.h, v_num, path.encode(), int(text), int(weighted),
                                        int(make_undirected), int(sort_adj)))

    def graph_from_csr(self, row_ptr, col, w=None):
        row_ptr = np.ascontiguousarray(row_ptr, np.uint64)
        col = np.ascontiguousarray(col, np.uint32)
        w = None if w is None else np.ascontiguousarray(w, np.float32)
        _check(_lib.kkw_graph_from_csr(self.h, len(row_ptr) - 1, _ptr(row_ptr), _ptr(col), _ptr(w)))

    def set_edge_types(self, etype):
        etype = np.ascontiguousarray(etype, np.uint8)
        _check(_lib.kkw_graph_set_edge_types(self.h, _ptr(etype)))

    def type_by_dst(self, ntypes, seed):
        _check(_lib.kkw_graph_type_by_dst(self.h, ntypes, seed))

    def info(self):
        v, e = C.c_uint32(), C.c_uint64()
        _check(_lib.kkw_graph_info(self.h, C.byref(v), C.byref(e)))
        return v.value, e.value

    def download(self, weights=True, etype=False):
        V, E = self.info()
        rp = np.zeros(V + 1, np.uint64)
        col = np.zeros(E, np.uint32)
        w = np.zeros(E, np.float32) if weights else None
        et = np.zeros(E, np.uint8) if etype else None
        _check(_lib.kkw_graph_download(self.h, _ptr(rp), _ptr(col), _ptr(w), _ptr(et)))
        return rp, col, w, et

    def prepare(self):
        _check(_lib.kkw_prepare(self.h))

    def adj_index(self, enable: bool = True):
        """Build (True) or drop (False) the neighbour hash sets of node2vec's adjacency test."""
        _check(_lib.kkw_graph_adj_index(self.h, 1 if enable else 0))

    def alias_download(self):
        V, E = self.info()
        p = np.zeros(E, np.float32)
        prim = np.zeros(E, np.uint32)
        al = np.zeros(E, np.uint32)
        meta = np.zeros(E, np.uint32)
        vs = np.zeros(V, np.float32)
        _check(_lib.kkw_alias_download(self.h, _ptr(p), _ptr(prim), _ptr(al), _ptr(meta), _ptr(vs)))
        return p, prim, al, meta, vs

    # ---- walking ---------------------------------------------------------
    def run(self, jobs: Sequence[Job], mode: int = MODE_PLAIN, stats: bool = False) -> List[Paths]:
        cj = (_Job * len(jobs))(*[j.to_c() for j in jobs])
        outs = (_Paths * len(jobs))()
        _check(_lib.kkw_run(self.h, cj, len(jobs), mode, OUT_HOST, FLAG_STATS if stats else 0, outs))
        return [_take_paths(outs[i]) for i in range(len(jobs))]

    def run_noalloc(self, jobs: Sequence[Job], mode: int, paths: Sequence[int], strides: Sequence[int],
                    lengths: Sequence[int]) -> None:
        """kkw_run_noalloc: walk into caller-owned device buffers (device
        addresses as integers: job i's rows at paths[i], strides[i] entries
        apart, its length words at lengths[i]); allocates nothing on the device."""
        n = len(jobs)
        cj = (_Job * n)(*[j.to_c() for j in jobs])
        pp = (C.c_void_p * n)(*paths)
        ss = (C.c_uint32 * n)(*strides)
        ll = (C.c_void_p * n)(*lengths)
        _check(_lib.kkw_run_noalloc(self.h, cj, n, mode, pp, ss, ll))

    def run_epochs(self, job: Job, epoch_walkers: int, consumer, mode: int = MODE_PERSISTENT) -> None:
        """kkw_run_epochs: walk `job` in epochs of epoch_walkers; consumer(paths,
        stream) gets each epoch's device paths (a _Paths with on_device = 1)
        and a HIP stream (int address) ordered after that epoch's walk, while
        the next epoch walks.  A consumer that returns False or raises stops
        the run (the exception is raised again here)."""
        errors = []

        def tramp(pp, stream, user):
            try:
                return 1 if consumer(pp.contents, stream) is False else 0
            except BaseException as e:   # re-raised below, after the library has returned
                errors.append(e)
                return 1

        fn = EPOCH_CONSUMER(tramp)
        cj = job.to_c()
        rc = _lib.kkw_run_epochs(self.h, C.byref(cj), epoch_walkers, mode, fn, None)
        if errors:
            raise errors[0]
        _check(rc)

    @staticmethod
    def download_paths(device_paths: "_Paths") -> Paths:
        """Host copy of device paths (kkw_paths_download)."""
        out = _Paths()
        _check(_lib.kkw_paths_download(C.byref(device_paths), C.byref(out)))
        return _take_paths(out)

    def batch(self, jobs: Sequence[Job], mode: int = MODE_PERSISTENT, flags: int = 0) -> Batch:
        return Batch(self, jobs, mode, flags)

    def program(self, source: str, options: Optional[str] = None) -> Program:
        """Compile a walk program (source text) for this engine's GPU."""
        return Program(self, source, options)

    def timing(self) -> dict:
        t = _Timing()
        _check(_lib.kkw_last_timing(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in _Timing._fields_}

    def stream_copy_ms(self, nbytes: int, iters: int = 10) -> float:
        """Average ms of a plain copy of nbytes (read + write) on this engine's stream."""
        ms = C.c_double()
        _check(_lib.kkw_diag_stream_copy(self.h, nbytes, iters, C.byref(ms)))
        return ms.value

    def random_gather_lines_per_s(self, nbytes: int, reads: int = 64, iters: int = 5) -> float:
        """Independent random 16-byte reads served per second, one 128-byte line each."""
        r = C.c_double()
        _check(_lib.kkw_diag_random_gather(self.h, nbytes, reads, iters, C.byref(r)))
        return r.value

    # ---- serving ---------------------------------------------------------
    def serve_start(self, max_walkers_in_flight=1 << 22, queue_capacity=1 << 16, round_steps=0, idle_poll_us=0):
        """Start the serving loop; round_steps > 0 walks each flight in kernel
        launches of at most that many steps per walker (same paths);
        idle_poll_us > 0 keeps the idle loop polling for new queries that long
        before it sleeps."""
        o = _ServeOpts(max_walkers_in_flight, round_steps, queue_capacity, idle_poll_us)
        _check(_lib.kkw_serve_start(self.h, C.byref(o)))

    def submit(self, job: Job) -> int:
        t = C.c_uint64()
        cj = job.to_c()
        _check(_lib.kkw_submit(self.h, C.byref(cj), C.byref(t)))
        return t.value

    def submit_c(self, cj: "_Job") -> int:
        """submit() of a descriptor already built with Job.to_c() (the Job
        must stay alive: the descriptor points into its arrays)."""
        t = C.c_uint64()
        _check(_lib.kkw_submit(self.h, C.byref(cj), C.byref(t)))
        return t.value

    def poll(self, ticket: int) -> bool:
        d = C.c_int()
        _check(_lib.kkw_poll(self.h, ticket, C.byref(d)))
        return bool(d.value)

    def wait(self, ticket: int, timeout_s: float = -1.0, packed: bool = False):
        """Result of a served query and its latency (ms, submit to completion).
        packed=True keeps the packed form (PackedPaths: one flat vertex array
        and offsets) instead of expanding rows to `stride` entries."""
        p = _Paths()
        lat = C.c_double()
        _check(_lib.kkw_wait(self.h, ticket, timeout_s, C.byref(p), C.byref(lat)))
        return (_take_packed(p) if packed else _take_paths(p)), lat.value

    def wait_steps(self, ticket: int, timeout_s: float = -1.0):
        """(steps, latency ms) of a served query; its paths are freed unread
        (load generators that only measure)."""
        p = _Paths()
        lat = C.c_double()
        _check(_lib.kkw_wait(self.h, ticket, timeout_s, C.byref(p), C.byref(lat)))
        steps = int(p.steps)
        _lib.kkw_paths_free(C.byref(p))
        return steps, lat.value

    def cancel(self, ticket: int):
        _check(_lib.kkw_cancel(self.h, ticket))

    def serve_stop(self):
        _check(_lib.kkw_serve_stop(self.h))
