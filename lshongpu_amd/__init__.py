"""lshongpu_amd -- MI355X-native k-NN search engine (Python side of the C ABI).

The compute path is liblshgpu.so (hand-written HIP for gfx950, built by
lshongpu_amd.build); this module only loads it with ctypes and mirrors the
reference's dispatch layer (LinearScan/kernel.cu:27-117):

    linearScans(implementation, k, d, N_query, N_data, data, queries,
                sketchDim, distanceFunc) -> Result
    LSH(implementation, keysImplementation, k, d, N_query, N_data, data,
        queries, sketchDim, distanceFunc, bucketKeyBits, tables,
        runWithSketchedData) -> Result

with the same argument meanings.  Failures raise LshError (the reference
prints and exits).  Index is the resident index for batched search and
serving.  There is no CPU fallback: if the library is missing, importing the
compute entry points raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

__all__ = [
    "LshError", "Result", "Index", "Server", "POINT_DTYPE", "load", "calculateK", "linearScans", "LSH",
    "linear_scans", "lsh_framework", "random_normals", "random_signs", "simhash_sketch",
    "jl_sketch", "gj_transform", "hamming_topc", "wmh_sketch", "crosspoly_keys", "crosspoly_dense_keys",
    "sort_table_keys", "autotune", "library_path",
    "read_text_dataset", "read_fvecs", "write_text_dataset", "read_truth", "write_truth", "validate",
    "write_results", "write_performance",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblshgpu.so")
_lib: Optional[C.CDLL] = None

# point.h:3-7 {int ID; float distance;}
POINT_DTYPE = np.dtype([("ID", np.int32), ("distance", np.float32)])


class LshError(RuntimeError):
    """A nonzero status from the C ABI."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


class _Point(C.Structure):
    _fields_ = [("ID", C.c_int32), ("distance", C.c_float)]


class _Result(C.Structure):
    # resultDTO.h:5-32 on LP64
    _fields_ = [
        ("results", C.POINTER(_Point)),
        ("preprocessTime", C.c_long),
        ("constructionTime", C.c_long),
        ("scanTime", C.c_long),
        ("recall", C.c_float),
        ("avgDistance", C.c_float),
    ]


class _Options(C.Structure):
    _fields_ = [
        ("candidates", C.c_int),
        ("device", C.c_int),
        ("rescan_margin", C.c_int),
        ("reserved", C.c_int * 8),
    ]


class _IndexConfig(C.Structure):
    _fields_ = [
        ("implementation", C.c_int),
        ("distanceFunc", C.c_int),
        ("d", C.c_int),
        ("N_data", C.c_int64),
        ("sketchDim", C.c_int),
        ("framework", C.c_int),
        ("keysImplementation", C.c_int),
        ("bucketKeyBits", C.c_int),
        ("tables", C.c_int),
        ("runWithSketchedData", C.c_int),
        ("candidates", C.c_int),
        ("device", C.c_int),
        ("rescan_margin", C.c_int),
        ("projection", C.c_int),
        ("reserved", C.c_int * 7),
    ]


_FP = C.POINTER(C.c_float)


class _ServerConfig(C.Structure):
    _fields_ = [("max_batch_queries", C.c_int), ("max_query_size", C.c_int), ("memory_budget_bytes", C.c_int64),
                ("reserved", C.c_int * 8)]


class _TuneReport(C.Structure):
    _fields_ = [("recall", C.c_double), ("search_ms", C.c_double), ("build_ms", C.c_double),
                ("device_bytes", C.c_int64), ("status", C.c_int)]


class _BucketStats(C.Structure):
    _fields_ = [("buckets", C.c_int64), ("nonempty", C.c_int64), ("points", C.c_int64), ("max_size", C.c_int64)]


_SIGS = {
    "lsh_calculate_k": (C.c_int, [C.c_int]),
    "lsh_options_default": (None, [C.POINTER(_Options)]),
    "lsh_linear_scans": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _FP, _FP, C.c_int, C.c_int,
                                   C.POINTER(_Result)]),
    "lsh_linear_scans_ex": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _FP, _FP, C.c_int, C.c_int,
                                      C.POINTER(_Options), C.POINTER(_Result)]),
    "lsh_framework": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _FP, _FP, C.c_int,
                                C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_Result)]),
    "lsh_framework_ex": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _FP, _FP, C.c_int,
                                   C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(_Options),
                                   C.POINTER(_Result)]),
    "lsh_result_free": (None, [C.POINTER(_Result)]),
    "lsh_last_error": (C.c_char_p, []),
    "lsh_version": (C.c_char_p, []),
    "lsh_random_normals": (C.c_int, [_FP, C.c_int64, C.c_double]),
    "lsh_random_signs": (C.c_int, [C.POINTER(C.c_int32), C.c_int64]),
    "lsh_simhash_sketch": (C.c_int, [_FP, C.c_int64, C.c_int, C.c_int, C.c_int, _FP, C.POINTER(C.c_uint32)]),
    "lsh_jl_sketch": (C.c_int, [_FP, C.c_int64, C.c_int, C.c_int, _FP, _FP]),
    "lsh_hamming_topc_noalloc": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int, C.c_int,
                                           C.c_void_p, C.c_void_p, C.c_void_p]),
    "lsh_gj_transform": (C.c_int, [_FP, C.c_int64, _FP, C.c_int64, C.c_int]),
    "lsh_hamming_topc": (C.c_int, [C.POINTER(C.c_uint32), C.c_int64, C.POINTER(C.c_uint32), C.c_int64, C.c_int,
                                   C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "lsh_wmh_sketch": (C.c_int, [_FP, C.c_int64, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint32)]),
    "lsh_crosspoly_keys": (C.c_int, [_FP, C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_int32),
                                     C.POINTER(C.c_uint16)]),
    "lsh_crosspoly_dense_keys": (C.c_int, [_FP, C.c_int64, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint32)]),
    "lsh_sort_table_keys": (C.c_int, [C.POINTER(C.c_uint32), C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_uint32),
                                      C.POINTER(C.c_uint32)]),
    "lsh_bucket_distribution": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]),
    "lsh_bucket_distribution_full_key": (C.c_int, [C.c_void_p, C.c_int64, C.c_int, C.POINTER(C.c_void_p),
                                                   C.POINTER(C.c_void_p), C.POINTER(C.c_int64)]),
    "lsh_free": (None, [C.c_void_p]),
    "lsh_read_text_dataset": (C.c_int, [C.c_char_p, C.POINTER(_FP), C.POINTER(C.c_int64), C.POINTER(C.c_int),
                                        C.POINTER(C.POINTER(C.c_int32))]),
    "lsh_read_fvecs": (C.c_int, [C.c_char_p, C.POINTER(_FP), C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "lsh_write_text_dataset": (C.c_int, [C.c_char_p, _FP, C.c_int64, C.c_int]),
    "lsh_read_truth": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                 C.POINTER(C.POINTER(C.c_int32)), C.POINTER(_FP)]),
    "lsh_write_truth": (C.c_int, [C.c_char_p, C.c_void_p, C.POINTER(C.c_int32), C.c_int, C.c_int, C.c_int]),
    "lsh_validate": (C.c_int, [C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float),
                               C.POINTER(C.c_float)]),
    "lsh_write_results": (C.c_int, [C.c_char_p, C.c_void_p, C.c_int, C.c_int, C.c_int]),
    "lsh_write_performance": (C.c_int, [C.c_char_p, C.POINTER(_Result), C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int]),
    "lsh_index_config_default": (None, [C.POINTER(_IndexConfig)]),
    "lsh_index_create": (C.c_int, [C.POINTER(_IndexConfig), C.c_void_p, C.c_int, C.POINTER(C.c_void_p)]),
    "lsh_index_search": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                   C.c_void_p]),
    "lsh_index_search_batches": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.c_int, C.c_int,
                                           C.c_int, C.c_void_p, C.c_int, C.c_void_p]),
    "lsh_index_set_profiling": (C.c_int, [C.c_void_p, C.c_int]),
    "lsh_index_batch_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "lsh_index_phase_times": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "lsh_index_phase_times_n": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.c_int]),
    "lsh_index_bucket_stats": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(_BucketStats), C.POINTER(C.c_int64),
                                         C.c_int]),
    "lsh_index_device_bytes": (C.c_int64, [C.c_void_p]),
    "lsh_server_config_default": (None, [C.POINTER(_ServerConfig)]),
    "lsh_server_create": (C.c_int, [C.c_void_p, C.POINTER(_ServerConfig), C.POINTER(C.c_void_p)]),
    "lsh_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.POINTER(C.c_int64)]),
    "lsh_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double)]),
    "lsh_server_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "lsh_server_destroy": (C.c_int, [C.c_void_p]),
    "lsh_autotune": (C.c_int, [C.POINTER(_IndexConfig), C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int,
                               C.c_double, C.POINTER(_TuneReport), C.POINTER(C.c_int)]),
    "lsh_index_destroy": (C.c_int, [C.c_void_p]),
}

EXPORTED_SYMBOLS = tuple(_SIGS)


def library_path() -> str:
    return _LIB_PATH


def load() -> C.CDLL:
    """Loads liblshgpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise LshError(-1, f"{_LIB_PATH} is missing: build it with `python -m lshongpu_amd.build`")
    # torch ships its own HIP runtime under the same soname; only one can be
    # loaded per process, and torch only works with its own, so it loads first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(_LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(status: int) -> None:
    if status != 0:
        msg = load().lsh_last_error().decode(errors="replace")
        raise LshError(status, msg)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(_FP)


@dataclass
class Result:
    """resultDTO.h:5-32: results is a (N_query, k) array of POINT_DTYPE."""

    results: np.ndarray
    preprocessTime: int
    constructionTime: int
    scanTime: int
    recall: float
    avgDistance: float

    @property
    def ids(self) -> np.ndarray:
        return self.results["ID"]

    @property
    def distances(self) -> np.ndarray:
        return self.results["distance"]


def calculateK(k: int) -> int:
    """launchHelper.cuh:21-25: k rounded up to a multiple of 32."""
    return int(load().lsh_calculate_k(int(k)))


def _options(candidates=None, device=0, rescan_margin=None) -> _Options:
    o = _Options()
    load().lsh_options_default(C.byref(o))
    if candidates is not None:
        o.candidates = int(candidates)
    o.device = int(device)
    if rescan_margin is not None:
        o.rescan_margin = int(rescan_margin)
    return o


def _take_result(r: _Result, n_query: int, k: int) -> Result:
    kq = calculateK(k)
    n = n_query * kq
    if n > 0:
        buf = np.ctypeslib.as_array(C.cast(r.results, C.POINTER(C.c_byte)), shape=(n * 8,)).copy()
        pts = buf.view(POINT_DTYPE).reshape(n_query, kq)
    else:
        pts = np.zeros((n_query, kq), dtype=POINT_DTYPE)
    res = Result(pts, int(r.preprocessTime), int(r.constructionTime), int(r.scanTime), float(r.recall),
                 float(r.avgDistance))
    load().lsh_result_free(C.byref(r))
    return res


def linearScans(implementation, k, d, N_query, N_data, data, queries, sketchDim, distanceFunc, *,
                candidates=None, device=0, rescan_margin=None) -> Result:
    """kernel.cu:27 linearScans (see include/lshgpu.h for the implementation ids)."""
    lib = load()
    data = _f32(data).reshape(-1)
    queries = _f32(queries).reshape(-1)
    if data.size < int(N_data) * int(d) or queries.size < int(N_query) * int(d):
        raise LshError(1, "data/queries are smaller than N x d")
    r = _Result()
    o = _options(candidates, device, rescan_margin)
    _check(lib.lsh_linear_scans_ex(int(implementation), int(k), int(d), int(N_query), int(N_data), _fp(data),
                                   _fp(queries), int(sketchDim), int(distanceFunc), C.byref(o), C.byref(r)))
    return _take_result(r, int(N_query), int(k))


def LSH(implementation, keysImplementation, k, d, N_query, N_data, data, queries, sketchDim, distanceFunc,
        bucketKeyBits, tables, runWithSketchedData, *, candidates=None, device=0, rescan_margin=None) -> Result:
    """kernel.cu:91 LSH: the (K,L) bucket framework."""
    lib = load()
    data = _f32(data).reshape(-1)
    queries = _f32(queries).reshape(-1)
    if data.size < int(N_data) * int(d) or queries.size < int(N_query) * int(d):
        raise LshError(1, "data/queries are smaller than N x d")
    r = _Result()
    o = _options(candidates, device, rescan_margin)
    _check(lib.lsh_framework_ex(int(implementation), int(keysImplementation), int(k), int(d), int(N_query),
                                int(N_data), _fp(data), _fp(queries), int(sketchDim), int(distanceFunc),
                                int(bucketKeyBits), int(tables), int(bool(runWithSketchedData)), C.byref(o),
                                C.byref(r)))
    return _take_result(r, int(N_query), int(k))


linear_scans = linearScans
lsh_framework = LSH


# ------------------------------------------------------------- stage entry points

def random_normals(n: int, stddev: float = 1.0) -> np.ndarray:
    out = np.empty(int(n), dtype=np.float32)
    _check(load().lsh_random_normals(_fp(out), int(n), float(stddev)))
    return out


def random_signs(n: int) -> np.ndarray:
    out = np.empty(int(n), dtype=np.int32)
    _check(load().lsh_random_signs(out.ctypes.data_as(C.POINTER(C.c_int32)), int(n)))
    return out


def simhash_sketch(x, words: int, hash_bits: int = 32, planes=None) -> np.ndarray:
    x = _f32(x)
    n, d = x.shape
    out = np.empty((n, int(words)), dtype=np.uint32)
    pp = None
    if planes is not None:
        planes = _f32(planes)
        pp = _fp(planes)
    _check(load().lsh_simhash_sketch(_fp(x), n, d, int(words), int(hash_bits), pp,
                                     out.ctypes.data_as(C.POINTER(C.c_uint32))))
    return out


def jl_sketch(x, sketch_dim: int, planes=None) -> np.ndarray:
    x = _f32(x)
    n, d = x.shape
    out = np.empty((n, int(sketch_dim)), dtype=np.float32)
    pp = None
    if planes is not None:
        planes = _f32(planes)
        pp = _fp(planes)
    _check(load().lsh_jl_sketch(_fp(x), n, d, int(sketch_dim), pp, _fp(out)))
    return out


def gj_transform(data, queries):
    data = _f32(data).copy()
    queries = _f32(queries).copy()
    d = data.shape[1]
    _check(load().lsh_gj_transform(_fp(data), data.shape[0], _fp(queries), queries.shape[0], d))
    return data, queries


def hamming_topc(data_sketch, query_sketch, C_: int):
    ds = np.ascontiguousarray(data_sketch, dtype=np.uint32)
    qs = np.ascontiguousarray(query_sketch, dtype=np.uint32)
    words = ds.shape[1]
    ids = np.empty((qs.shape[0], int(C_)), dtype=np.int32)
    dist = np.empty((qs.shape[0], int(C_)), dtype=np.int32)
    _check(load().lsh_hamming_topc(ds.ctypes.data_as(C.POINTER(C.c_uint32)), ds.shape[0],
                                   qs.ctypes.data_as(C.POINTER(C.c_uint32)), qs.shape[0], words, int(C_),
                                   ids.ctypes.data_as(C.POINTER(C.c_int32)),
                                   dist.ctypes.data_as(C.POINTER(C.c_int32))))
    return ids, dist


def wmh_sketch(x, sketch_dim: int, one_bit: bool = False) -> np.ndarray:
    """Weighted MinHash of rows that are already shifted and normalised."""
    x = _f32(x)
    n, d = x.shape
    out = np.empty((n, int(sketch_dim)), dtype=np.uint32)
    _check(load().lsh_wmh_sketch(_fp(x), n, d, int(sketch_dim), int(bool(one_bit)),
                                 out.ctypes.data_as(C.POINTER(C.c_uint32))))
    return out


def crosspoly_keys(x, tables: int, signs=None) -> np.ndarray:
    x = _f32(x)
    n, d = x.shape
    keys = np.empty((n, int(tables)), dtype=np.uint16)
    sp = None
    if signs is not None:
        signs = np.ascontiguousarray(signs, dtype=np.int32)
        sp = signs.ctypes.data_as(C.POINTER(C.c_int32))
    _check(load().lsh_crosspoly_keys(_fp(x), n, d, int(tables), sp, keys.ctypes.data_as(C.POINTER(C.c_uint16))))
    return keys


def crosspoly_dense_keys(x, tables: int, K: int) -> np.ndarray:
    """keysImplementation 8 keys [n][tables] (uint32) of the rows of x (normalised first)."""
    x = _f32(x)
    n, d = x.shape
    keys = np.empty((n, tables), dtype=np.uint32)
    _check(load().lsh_crosspoly_dense_keys(_fp(x), n, d, int(tables), int(K),
                                           keys.ctypes.data_as(C.POINTER(C.c_uint32))))
    return keys


def sort_table_keys(keys, bits: int):
    """Stable device sort of every row of keys [tables][N] (low `bits` bits):
    (sorted keys, IDs), both [tables][N]."""
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    L, N = keys.shape
    sk = np.empty_like(keys)
    ids = np.empty_like(keys)
    u32 = C.POINTER(C.c_uint32)
    _check(load().lsh_sort_table_keys(keys.ctypes.data_as(u32), N, L, int(bits), sk.ctypes.data_as(u32),
                                      ids.ctypes.data_as(u32)))
    return sk, ids


def bucket_distribution(hashes) -> np.ndarray:
    """statistics.cuh:10-16 bucketDistribution: counts of each byte value 0..255."""
    h = np.ascontiguousarray(hashes, dtype=np.uint8).reshape(-1)
    counts = np.zeros(256, dtype=np.int64)
    _check(load().lsh_bucket_distribution(C.c_void_p(h.ctypes.data), h.size,
                                          counts.ctypes.data_as(C.POINTER(C.c_int64))))
    return counts


def bucket_distribution_full_key(hashes, sketchDim: int | None = None):
    """statisticsCpu.cpp:11-28 bucketDistributionFullKey on (n, sketchDim) uint8
    keys: (distinct keys in byte-lexicographic order, their counts)."""
    h = np.ascontiguousarray(hashes, dtype=np.uint8)
    if sketchDim is None:
        if h.ndim != 2:
            raise LshError(1, "hashes must be (n, sketchDim) or sketchDim given")
        sketchDim = h.shape[1]
    sd = int(sketchDim)
    if sd < 1 or h.size % sd:
        raise LshError(1, "hashes do not hold whole keys of sketchDim bytes")
    n = h.size // sd
    lib = load()
    kp, cp, nd = C.c_void_p(), C.c_void_p(), C.c_int64()
    _check(lib.lsh_bucket_distribution_full_key(C.c_void_p(h.ctypes.data), n, sd, C.byref(kp), C.byref(cp),
                                                C.byref(nd)))
    try:
        m = int(nd.value)
        keys = np.ctypeslib.as_array(C.cast(kp, C.POINTER(C.c_uint8)), shape=(m * sd,)).copy().reshape(m, sd) \
            if m else np.zeros((0, sd), np.uint8)
        counts = np.ctypeslib.as_array(C.cast(cp, C.POINTER(C.c_int64)), shape=(m,)).copy() \
            if m else np.zeros(0, np.int64)
    finally:
        lib.lsh_free(kp)
        lib.lsh_free(cp)
    return keys, counts


# ------------------------------------------------------------------ resident index

def _fill_config(cfg, n, d, *, implementation=2, distanceFunc=1, sketchDim=0, framework=0, keysImplementation=0,
                 bucketKeyBits=0, tables=0, runWithSketchedData=False, candidates=None, device=0,
                 rescan_margin=None, projection=0) -> None:
    cfg.implementation = int(implementation)
    cfg.distanceFunc = int(distanceFunc)
    cfg.d = int(d)
    cfg.N_data = int(n)
    cfg.sketchDim = int(sketchDim)
    cfg.framework = int(framework)
    cfg.keysImplementation = int(keysImplementation)
    cfg.bucketKeyBits = int(bucketKeyBits)
    cfg.tables = int(tables)
    cfg.runWithSketchedData = int(bool(runWithSketchedData))
    if candidates is not None:
        cfg.candidates = int(candidates)
    if rescan_margin is not None:
        cfg.rescan_margin = int(rescan_margin)
    cfg.device = int(device)
    cfg.projection = int(projection)


def autotune(data, queries, k: int, candidates, target_recall: float = 0.9):
    """Measures each candidate configuration (a list of Index keyword dicts)
    on `data` (host array or device tensor) with the host query sample, and
    returns (index of the chosen one, list of reports: recall, search_ms,
    build_ms, device_bytes, status).  The fastest candidate reaching
    target_recall is chosen (lsh_autotune)."""
    lib = load()
    queries = _f32(queries)
    if hasattr(data, "data_ptr"):
        _check_tensor(data, "data", "float32")
        n, d = data.shape
    else:
        data = _f32(data)
        n, d = data.shape
    arr = (_IndexConfig * len(candidates))()
    for i, kw in enumerate(candidates):
        lib.lsh_index_config_default(C.byref(arr[i]))
        _fill_config(arr[i], n, d, **kw)
    reps = (_TuneReport * len(candidates))()
    chosen = C.c_int(-1)
    dp, on_dev = _ptr_of(data)
    _check(lib.lsh_autotune(arr, len(candidates), dp, int(on_dev), C.c_void_p(queries.ctypes.data),
                            int(queries.shape[0]), int(k), float(target_recall), reps, C.byref(chosen)))
    return int(chosen.value), [{f: getattr(r, f) for f, _ in _TuneReport._fields_} for r in reps]


def _check_tensor(t, name: str, dtype) -> None:
    """A torch tensor handed to the library as a raw pointer must be what the
    kernels assume: contiguous, of the right dtype, and (for 2-D operands) 2-D."""
    if not t.is_contiguous():
        raise LshError(1, f"{name} tensor must be contiguous")
    if dtype is not None:
        if str(t.dtype) != f"torch.{dtype}":
            raise LshError(1, f"{name} tensor must be {dtype}, not {str(t.dtype).replace('torch.', '')}")
        if t.dim() != 2:
            raise LshError(1, f"{name} tensor must be 2-D, not {t.dim()}-D")


def _ptr_of(a):
    """(pointer, on_device) of a numpy array or a torch tensor."""
    if hasattr(a, "data_ptr") and hasattr(a, "is_cuda"):
        return C.c_void_p(a.data_ptr()), bool(a.is_cuda)
    return C.c_void_p(a.ctypes.data), False


class Index:
    """Resident index: data, sketches and tables stay in HBM between searches.

    data may be a float32 numpy array (host) or a contiguous float32 torch
    tensor on the device.
    """

    def __init__(self, data, *, implementation=2, distanceFunc=1, sketchDim=0, framework=0,
                 keysImplementation=0, bucketKeyBits=0, tables=0, runWithSketchedData=False, candidates=None,
                 device=0, rescan_margin=None, projection=0):
        lib = load()
        cfg = _IndexConfig()
        lib.lsh_index_config_default(C.byref(cfg))
        if hasattr(data, "data_ptr"):
            _check_tensor(data, "data", "float32")
            n, d = data.shape
        else:
            data = _f32(data)
            if data.ndim != 2:
                raise LshError(1, f"data must be a 2-D (N_data, d) array, not {data.ndim}-D")
            n, d = data.shape
        _fill_config(cfg, n, d, implementation=implementation, distanceFunc=distanceFunc, sketchDim=sketchDim,
                     framework=framework, keysImplementation=keysImplementation, bucketKeyBits=bucketKeyBits,
                     tables=tables, runWithSketchedData=runWithSketchedData, candidates=candidates, device=device,
                     rescan_margin=rescan_margin, projection=projection)
        self.d = int(d)
        self.n = int(n)
        self._h = C.c_void_p()
        ptr, on_dev = _ptr_of(data)
        _check(lib.lsh_index_create(C.byref(cfg), ptr, int(on_dev), C.byref(self._h)))

    def search(self, queries, k: int, out=None, stream=None):
        """Returns (N_query, calculateK(k)) POINT_DTYPE results (numpy), or
        fills `out` (a device int64/float tensor view of the same bytes)."""
        lib = load()
        if int(k) < 1:
            raise LshError(1, "k must be >= 1")
        kq = calculateK(k)
        if hasattr(queries, "data_ptr"):
            _check_tensor(queries, "queries", "float32")
        else:
            queries = _f32(queries)
            if queries.ndim != 2:
                raise LshError(1, f"queries must be a 2-D (N_query, d) array, not {queries.ndim}-D")
        if int(queries.shape[1]) != self.d:
            raise LshError(1, f"queries have {int(queries.shape[1])} columns; the index holds d = {self.d}")
        nq = int(queries.shape[0])
        qp, q_dev = _ptr_of(queries)
        if out is None:
            res = np.empty((nq, kq), dtype=POINT_DTYPE)
            rp, r_dev = C.c_void_p(res.ctypes.data), False
        else:
            need = nq * kq * POINT_DTYPE.itemsize
            if hasattr(out, "data_ptr"):
                _check_tensor(out, "out", None)
                have = out.numel() * out.element_size()
            else:
                if not (isinstance(out, np.ndarray) and out.flags.c_contiguous and out.flags.writeable):
                    raise LshError(1, "out must be a C-contiguous, writeable numpy array or device tensor")
                have = out.nbytes
            if have < need:
                raise LshError(1, f"out holds {have} bytes; {nq} queries x {kq} results need {need}")
            res = out
            rp, r_dev = _ptr_of(out)
        sp = C.c_void_p(stream) if stream else None
        _check(lib.lsh_index_search(self._h, qp, nq, int(k), int(q_dev), rp, int(r_dev), sp))
        return res

    def search_batches(self, queries, batch_sizes, k: int, out=None, stream=None):
        """Each batch (consecutive rows of queries, batch_sizes[i] of them) as
        its own search, through captured graphs where the index supports
        them (lsh_index_search_batches); results as search() for all rows."""
        lib = load()
        if int(k) < 1:
            raise LshError(1, "k must be >= 1")
        kq = calculateK(k)
        if hasattr(queries, "data_ptr"):
            _check_tensor(queries, "queries", "float32")
        else:
            queries = _f32(queries)
            if queries.ndim != 2:
                raise LshError(1, "queries must be a 2-D (N_query, d) array")
        if int(queries.shape[1]) != self.d:
            raise LshError(1, f"queries have {int(queries.shape[1])} columns; the index holds d = {self.d}")
        sizes = np.ascontiguousarray(batch_sizes, dtype=np.int32)
        nq = int(queries.shape[0])
        if sizes.ndim != 1 or (sizes < 0).any() or int(sizes.sum()) != nq:
            raise LshError(1, "batch sizes must be non-negative and sum to the number of queries")
        if out is None:
            res = np.empty((nq, kq), dtype=POINT_DTYPE)
        else:
            have = out.numel() * out.element_size() if hasattr(out, "data_ptr") else out.nbytes
            if hasattr(out, "data_ptr"):
                _check_tensor(out, "out", None)
            if have < nq * kq * POINT_DTYPE.itemsize:
                raise LshError(1, "out is too small")
            res = out
        qp, q_dev = _ptr_of(queries)
        rp, r_dev = _ptr_of(res)
        sp = C.c_void_p(stream) if stream else None
        _check(lib.lsh_index_search_batches(self._h, qp, sizes.ctypes.data_as(C.POINTER(C.c_int)), len(sizes),
                                            int(k), int(q_dev), rp, int(r_dev), sp))
        return res

    def batch_stats(self) -> dict:
        """The last search_batches call: batches that replayed a captured
        graph and queries redone by ordinary searches (lsh_index_batch_stats)."""
        g, r = C.c_int64(), C.c_int64()
        _check(load().lsh_index_batch_stats(self._h, C.byref(g), C.byref(r)))
        return {"graph_batches": g.value, "redone_queries": r.value}

    def set_profiling(self, on=True) -> None:
        """True / 1: time every phase; 2: only the rerank / finalisation phase;
        3: only the pivot-cell search's appending pass; False / 0: off."""
        _check(load().lsh_index_set_profiling(self._h, int(on)))

    def phase_times(self) -> dict:
        """Device time per phase (ms) and matrix-core scan counters, summed over
        the searches since profiling was turned on or since the last call
        (which waits for those searches)."""
        arr = (C.c_double * 11)()
        _check(load().lsh_index_phase_times_n(self._h, arr, 11))
        v = list(arr)
        return {"prep_ms": v[0], "scan_ms": v[1], "select_ms": v[2], "final_ms": v[3], "total_ms": v[4],
                "mfma_scan_ms": v[5], "mfma_launches": int(v[6]), "mfma_points": int(v[7]), "recheck_ms": v[8],
                "pass2_ms": v[9], "pass2_tiles": int(v[10])}

    def bucket_stats(self, table: int, hist_len: int = 0):
        """Occupancy of one (K,L) table: ({buckets, nonempty, points, max_size},
        histogram of bucket sizes whose last bin collects the larger ones)."""
        st = _BucketStats()
        hist = np.zeros(max(int(hist_len), 0), dtype=np.int64)
        hp = hist.ctypes.data_as(C.POINTER(C.c_int64)) if hist.size else None
        _check(load().lsh_index_bucket_stats(self._h, int(table), C.byref(st), hp, int(hist.size)))
        return {f: int(getattr(st, f)) for f, _ in _BucketStats._fields_}, hist

    def device_bytes(self) -> int:
        return int(load().lsh_index_device_bytes(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            load().lsh_index_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Server:
    """Request loop over a resident Index (lsh_server_*): submit() queues a
    request of host queries and returns a ticket; wait(ticket) returns its
    (results, latency_ms).  Requests are batched continuously by one worker
    thread under a device memory budget; submit and wait may be called from
    any thread."""

    def __init__(self, index: "Index", max_batch_queries: int = 4096, max_query_size=None,
                 memory_budget_bytes: int = 0):
        if max_query_size is None:
            max_query_size = min(2048, int(max_batch_queries))
        lib = load()
        cfg = _ServerConfig()
        lib.lsh_server_config_default(C.byref(cfg))
        cfg.max_batch_queries = int(max_batch_queries)
        cfg.max_query_size = int(max_query_size)
        cfg.memory_budget_bytes = int(memory_budget_bytes)
        self._index = index  # kept alive while the server runs
        self._h = C.c_void_p()
        self._live = {}
        _check(lib.lsh_server_create(index._h, C.byref(cfg), C.byref(self._h)))

    def submit(self, queries, k: int) -> int:
        q = _f32(queries)
        if q.ndim != 2 or q.shape[1] != self._index.d:
            raise LshError(1, f"queries must be (n, {self._index.d})")
        res = np.empty((q.shape[0], calculateK(k)), dtype=POINT_DTYPE)
        t = C.c_int64()
        _check(load().lsh_server_submit(self._h, C.c_void_p(q.ctypes.data), int(q.shape[0]), int(k),
                                        C.c_void_p(res.ctypes.data), C.byref(t)))
        self._live[t.value] = (q, res)
        return t.value

    def wait(self, ticket: int):
        lat = C.c_double()
        st = load().lsh_server_wait(self._h, int(ticket), C.byref(lat))
        q, res = self._live.pop(int(ticket), (None, None))
        _check(st)
        return res, float(lat.value)

    def stats(self) -> dict:
        b, q, m = C.c_int64(), C.c_int64(), C.c_int()
        _check(load().lsh_server_stats(self._h, C.byref(b), C.byref(q), C.byref(m)))
        return {"batches": b.value, "queries": q.value, "max_batch": m.value}

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            load().lsh_server_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


# ------------------------------------------------------------------ runner I/O
# Datasets, truth files, validation and report formats of the reference's
# runner (include/lshgpu.h documents the formats and their reference lines).

def _path(p) -> bytes:
    return os.fsencode(p)


def _points(results) -> np.ndarray:
    """(N_query, k) POINT_DTYPE array, or a Result."""
    if isinstance(results, Result):
        results = results.results
    r = np.ascontiguousarray(results)
    if r.dtype != POINT_DTYPE or r.ndim != 2:
        raise LshError(1, "results must be a (N_query, k) array of POINT_DTYPE")
    return r


def read_text_dataset(path):
    """gloveparser.cu:9-86 format; returns (data float32 [N][d], ids int32 [N])."""
    lib = load()
    data, ids = _FP(), C.POINTER(C.c_int32)()
    n, d = C.c_int64(), C.c_int()
    _check(lib.lsh_read_text_dataset(_path(path), C.byref(data), C.byref(n), C.byref(d), C.byref(ids)))
    try:
        x = np.ctypeslib.as_array(data, shape=(max(n.value, 1) * d.value,))[: n.value * d.value].copy()
        i = np.ctypeslib.as_array(ids, shape=(max(n.value, 1),))[: n.value].copy()
    finally:
        lib.lsh_free(C.cast(data, C.c_void_p))
        lib.lsh_free(C.cast(ids, C.c_void_p))
    return x.reshape(n.value, d.value), i


def read_fvecs(path) -> np.ndarray:
    """siftParser.cpp:8-28 .fvecs records; returns float32 [N][d]."""
    lib = load()
    data = _FP()
    n, d = C.c_int64(), C.c_int()
    _check(lib.lsh_read_fvecs(_path(path), C.byref(data), C.byref(n), C.byref(d)))
    try:
        x = np.ctypeslib.as_array(data, shape=(n.value * d.value,)).copy()
    finally:
        lib.lsh_free(C.cast(data, C.c_void_p))
    return x.reshape(n.value, d.value)


def write_text_dataset(path, data) -> None:
    """writer.cpp:10-29 format (IDs 0..N-1)."""
    x = _f32(data)
    if x.ndim != 2:
        raise LshError(1, "data must be 2-D")
    _check(load().lsh_write_text_dataset(_path(path), _fp(x), x.shape[0], x.shape[1]))


def read_truth(path):
    """goldstandard.cpp:123-134 truth file; returns (ids int32 [n][k], distances float32 [n][k])."""
    lib = load()
    ids, dist = C.POINTER(C.c_int32)(), _FP()
    n, k = C.c_int(), C.c_int()
    _check(lib.lsh_read_truth(_path(path), C.byref(n), C.byref(k), C.byref(ids), C.byref(dist)))
    m = n.value * k.value
    try:
        i = np.ctypeslib.as_array(ids, shape=(max(m, 1),))[:m].copy()
        dd = np.ctypeslib.as_array(dist, shape=(max(m, 1),))[:m].copy()
    finally:
        lib.lsh_free(C.cast(ids, C.c_void_p))
        lib.lsh_free(C.cast(dist, C.c_void_p))
    return i.reshape(n.value, k.value), dd.reshape(n.value, k.value)


def write_truth(path, results, k: int, query_ids=None) -> None:
    """Writes the first k entries of every result row as a truth file."""
    r = _points(results)
    qid = None
    if query_ids is not None:
        qid = np.ascontiguousarray(query_ids, dtype=np.int32)
    _check(load().lsh_write_truth(_path(path), r.ctypes.data_as(C.c_void_p),
                                  qid.ctypes.data_as(C.POINTER(C.c_int32)) if qid is not None else None,
                                  r.shape[0], r.shape[1], int(k)))


def validate(truth_path, results, reportK: int):
    """validation.cpp:129-208 runValidation; returns (recall, avgDistance)."""
    r = _points(results)
    rec, avg = C.c_float(), C.c_float()
    _check(load().lsh_validate(_path(truth_path), r.ctypes.data_as(C.c_void_p), r.shape[0], r.shape[1],
                               int(reportK), C.byref(rec), C.byref(avg)))
    return float(rec.value), float(avg.value)


def write_results(path, results, reportK: int) -> None:
    """resultWriter.cpp:7-20 writeResult (path None: stdout)."""
    r = _points(results)
    _check(load().lsh_write_results(_path(path) if path is not None else None, r.ctypes.data_as(C.c_void_p),
                                    r.shape[1], r.shape[0], int(reportK)))


def write_performance(path, result: Result, implementation: int, keysImplementation: int, sketchDim: int, k: int,
                      bucketKeyBits: int, tables: int) -> None:
    """resultWriter.cpp:37-60 writePerformanceResults: appends a 13-column CSV line."""
    r = _Result()
    r.preprocessTime = int(result.preprocessTime)
    r.constructionTime = int(result.constructionTime)
    r.scanTime = int(result.scanTime)
    r.recall = float(result.recall)
    r.avgDistance = float(result.avgDistance)
    _check(load().lsh_write_performance(_path(path), C.byref(r), int(implementation), int(keysImplementation),
                                        int(sketchDim), int(k), int(bucketKeyBits), int(tables)))
