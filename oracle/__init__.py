"""CPU oracle of the k-NN path (ctypes bindings to oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never by the product.
See oracle.cpp for the reference file:line each function restates.

Parity pins of each function (what checks it independently of itself) are
listed in tests/test_oracle.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_LIB = os.path.join(REF_DIR, "libref.so")
REF_GOLDSTANDARD = os.path.join(REF_DIR, "goldstandard")
REFERENCE_ROOT = "/root/reference"

_lib = None
_FP = C.POINTER(C.c_float)
_I32 = C.POINTER(C.c_int32)
_U32 = C.POINTER(C.c_uint32)


def build(ref: bool = True) -> None:
    """Compiles liboracle.so, and oracle/_ref/ when the reference sources exist."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    if ref and os.path.isdir(REFERENCE_ROOT):
        subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={REFERENCE_ROOT}"], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build(ref=False)
    lib = C.CDLL(LIB)
    sig = {
        "orc_calculate_k": (C.c_int, [C.c_int]),
        "orc_dist64": (C.c_double, [C.c_int, _FP, _FP, C.c_int]),
        "orc_exact_knn": (None, [_FP, C.c_int64, _FP, C.c_int64, C.c_int, C.c_int, C.c_int, _I32, _FP]),
        "orc_normalize_rows": (None, [_FP, C.c_int64, C.c_int]),
        "orc_gj_transform": (None, [_FP, C.c_int64, _FP, C.c_int64, C.c_int]),
        "orc_normal_stream": (None, [_FP, C.c_int64, C.c_double]),
        "orc_sign_stream": (None, [_I32, C.c_int64]),
        "orc_bit_stream": (None, [C.POINTER(C.c_uint8), C.c_int64]),
        "orc_simhash": (None, [_FP, C.c_int64, C.c_int, _FP, C.c_int, C.c_int, _U32]),
        "orc_jl": (None, [_FP, C.c_int64, C.c_int, _FP, C.c_int, _FP]),
        "orc_hamming": (C.c_int, [_U32, _U32, C.c_int]),
        "orc_sketch_topc": (None, [C.c_int, _U32, C.c_int64, _U32, C.c_int64, C.c_int, C.c_int, _I32, _I32]),
        "orc_jl_topc": (None, [_FP, C.c_int64, _FP, C.c_int64, C.c_int, C.c_int, _I32]),
        "orc_crosspoly_dense": (None, [_FP, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, _FP, _U32]),
        "orc_rerank": (None, [_FP, _FP, C.c_int64, C.c_int, C.c_int, _I32, C.c_int, C.c_int, _I32, _FP]),
        "orc_recall": (C.c_float, [_I32, C.c_int, _I32, C.c_int, C.c_int64, C.c_int]),
        "orc_distance_ratio": (C.c_float, [_FP, C.c_int, _I32, _FP, C.c_int, C.c_int64, C.c_int]),
        "orc_pow2_at_least": (C.c_int, [C.c_int]),
        "orc_crosspoly": (None, [_FP, C.c_int64, C.c_int, C.c_int, _I32, C.POINTER(C.c_uint16)]),
        "orc_philox4x32_10": (None, [_U32, _U32, _U32]),
        "orc_wmh_uniform": (C.c_float, [C.c_uint32, C.c_uint32]),
        "orc_wmh_bounds": (C.c_int, [_FP, C.c_int64, _FP, C.c_int64, C.c_int, _I32, _I32]),
        "orc_wmh_sketch": (None, [_FP, C.c_int64, C.c_int, _I32, _I32, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.POINTER(C.c_uint8), C.c_int, _U32]),
        "orc_lsh_candidates": (C.c_int64, [_U32, C.c_int64, _U32, C.c_int, C.c_int64, _I32, C.c_int64]),
        "orc_goldstandard_port": (C.c_double, [_FP, C.c_int64, _FP, C.c_int64, C.c_int, C.c_int, C.c_int, _I32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_FP)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_I32)


def _u(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_U32)


def calculate_k(k: int) -> int:
    return int(load().orc_calculate_k(int(k)))


# Distance kinds: 0 = -cos (implementation 1), 1 angular, 2 generalized Jaccard, 3 Euclidean.
def dist64(func: int, q, x) -> float:
    q, qp = _f(q)
    x, xp = _f(x)
    return float(load().orc_dist64(int(func), qp, xp, q.shape[-1]))


def exact_knn(data, queries, k: int, func: int = 1):
    data, dp = _f(data)
    queries, qp = _f(queries)
    n, d = data.shape
    q = queries.shape[0]
    ids = np.empty((q, k), dtype=np.int32)
    dist = np.empty((q, k), dtype=np.float32)
    load().orc_exact_knn(dp, n, qp, q, d, int(func), int(k), ids.ctypes.data_as(_I32), dist.ctypes.data_as(_FP))
    return ids, dist


def normalize_rows(x):
    x = np.array(x, dtype=np.float32, copy=True, order="C")
    load().orc_normalize_rows(x.ctypes.data_as(_FP), x.shape[0], x.shape[1])
    return x


def gj_transform(data, queries):
    data = np.array(data, dtype=np.float32, copy=True, order="C")
    queries = np.array(queries, dtype=np.float32, copy=True, order="C")
    load().orc_gj_transform(data.ctypes.data_as(_FP), data.shape[0], queries.ctypes.data_as(_FP),
                            queries.shape[0], data.shape[1])
    return data, queries


def normal_stream(n: int, stddev: float = 1.0):
    out = np.empty(int(n), dtype=np.float32)
    load().orc_normal_stream(out.ctypes.data_as(_FP), int(n), float(stddev))
    return out


def sign_stream(n: int):
    out = np.empty(int(n), dtype=np.int32)
    load().orc_sign_stream(out.ctypes.data_as(_I32), int(n))
    return out


def bit_stream(n: int):
    out = np.empty(int(n), dtype=np.uint8)
    load().orc_bit_stream(out.ctypes.data_as(C.POINTER(C.c_uint8)), int(n))
    return out


def simhash(x, planes, words: int, hash_bits: int = 32):
    x, xp = _f(x)
    planes, pp = _f(planes)
    out = np.empty((x.shape[0], words), dtype=np.uint32)
    load().orc_simhash(xp, x.shape[0], x.shape[1], pp, int(words), int(hash_bits), out.ctypes.data_as(_U32))
    return out


def jl(x, planes, sdim: int):
    x, xp = _f(x)
    planes, pp = _f(planes)
    out = np.empty((x.shape[0], sdim), dtype=np.float32)
    load().orc_jl(xp, x.shape[0], x.shape[1], pp, int(sdim), out.ctypes.data_as(_FP))
    return out


def sketch_topc(kind: int, data_sketch, query_sketch, C_: int):
    ds, dp = _u(data_sketch)
    qs, qp = _u(query_sketch)
    ids = np.empty((qs.shape[0], C_), dtype=np.int32)
    dist = np.empty((qs.shape[0], C_), dtype=np.int32)
    load().orc_sketch_topc(int(kind), dp, ds.shape[0], qp, qs.shape[0], ds.shape[1], int(C_),
                           ids.ctypes.data_as(_I32), dist.ctypes.data_as(_I32))
    return ids, dist


def jl_topc(data_sketch, query_sketch, C_: int):
    ds, dp = _f(data_sketch)
    qs, qp = _f(query_sketch)
    ids = np.empty((qs.shape[0], C_), dtype=np.int32)
    load().orc_jl_topc(dp, ds.shape[0], qp, qs.shape[0], ds.shape[1], int(C_), ids.ctypes.data_as(_I32))
    return ids


def rerank(data, queries, cand, k: int, func: int):
    data, dp = _f(data)
    queries, qp = _f(queries)
    cand, cp = _i(cand)
    q = queries.shape[0]
    ids = np.empty((q, k), dtype=np.int32)
    dist = np.empty((q, k), dtype=np.float32)
    load().orc_rerank(dp, qp, q, data.shape[1], int(func), cp, cand.shape[1], int(k), ids.ctypes.data_as(_I32),
                      dist.ctypes.data_as(_FP))
    return ids, dist


def recall(truth_ids, res_ids, report_k: int) -> float:
    t, tp = _i(truth_ids)
    r, rp = _i(res_ids)
    return float(load().orc_recall(tp, t.shape[1], rp, r.shape[1], t.shape[0], int(report_k)))


def distance_ratio(truth_d, res_ids, res_d, report_k: int) -> float:
    t, tp = _f(truth_d)
    r, rp = _i(res_ids)
    rd, rdp = _f(res_d)
    return float(load().orc_distance_ratio(tp, t.shape[1], rp, rdp, r.shape[1], t.shape[0], int(report_k)))


def crosspoly(x_normalized, tables: int, signs):
    x, xp = _f(x_normalized)
    s, sp = _i(signs)
    keys = np.empty((x.shape[0], tables), dtype=np.uint16)
    load().orc_crosspoly(xp, x.shape[0], x.shape[1], int(tables), sp, keys.ctypes.data_as(C.POINTER(C.c_uint16)))
    return keys


def pow2_at_least(d: int) -> int:
    P = 1
    while P < d:
        P *= 2
    return P


def bf16_round(x):
    """float32 values rounded to the nearest bfloat16 (ties to even), as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return r.astype(np.uint32).view(np.float32)


def dense_rotations(tables: int, K: int, P: int, d: int):
    """[tables*K][P][d] Gaussian rotations of keysImplementation 8: the
    reference's default-seeded normal stream (randomVectorGenerator.cpp:32-53),
    each value rounded to bfloat16 so that it is exact on the matrix cores."""
    return bf16_round(normal_stream(tables * K * P * d)).reshape(tables * K, P, d)


def crosspoly_dense(x_normalized, tables: int, K: int, G=None):
    """Keys [n][tables] (uint32) of keysImplementation 8 (orc_crosspoly_dense)."""
    x, xp = _f(x_normalized)
    n, d = x.shape
    P = pow2_at_least(d)
    if G is None:
        G = dense_rotations(tables, K, P, d)
    G, gp = _f(G)
    assert G.shape == (tables * K, P, d)
    keys = np.empty((n, tables), dtype=np.uint32)
    load().orc_crosspoly_dense(xp, n, d, P, tables, K, gp, keys.ctypes.data_as(_U32))
    return keys


def philox4x32_10(ctr, key):
    c, cp = _u(ctr)
    k, kp = _u(key)
    out = np.empty(4, dtype=np.uint32)
    load().orc_philox4x32_10(cp, kp, out.ctypes.data_as(_U32))
    return out


def wmh_uniform(seed: int, t: int) -> float:
    return float(load().orc_wmh_uniform(int(seed) & 0xFFFFFFFF, int(t)))


def wmh_bounds(x, y=None):
    x, xp = _f(x)
    if y is None:
        y = np.zeros((0, x.shape[1]), dtype=np.float32)
    y, yp = _f(y)
    d = x.shape[1]
    bounds = np.empty(d, dtype=np.int32)
    start = np.empty(d, dtype=np.int32)
    M = load().orc_wmh_bounds(xp, x.shape[0], yp, y.shape[0], d, bounds.ctypes.data_as(_I32),
                              start.ctypes.data_as(_I32))
    return bounds, start, int(M)


def wmh_sketch(x, bounds, start, M: int, H: int, mode: int, bitmap, hbits: int = 32):
    """H counts (mode 0) or H words of `hbits` one-bit hashes (mode 1) per row."""
    x, xp = _f(x)
    b, bp = _i(bounds)
    s, sp = _i(start)
    bm = np.ascontiguousarray(bitmap, dtype=np.uint8)
    out = np.empty((x.shape[0], H), dtype=np.uint32)
    load().orc_wmh_sketch(xp, x.shape[0], x.shape[1], bp, sp, int(M), int(H), int(mode), int(hbits),
                          bm.ctypes.data_as(C.POINTER(C.c_uint8)), bm.shape[0], out.ctypes.data_as(_U32))
    return out


def lsh_candidates(dkeys, qkeys, q: int):
    dk, dp = _u(dkeys)
    qk, qp = _u(qkeys)
    n, tables = dk.shape
    out = np.empty(n, dtype=np.int32)
    c = load().orc_lsh_candidates(dp, n, qp, tables, int(q), out.ctypes.data_as(_I32), n)
    return out[:c].copy()


def lsh_keys(x, keys_impl: int, tables: int, bits: int, wmh=None):
    """Bucket keys [n][tables] of rows that have had the key family's
    preprocessing (lshFramework.cuh:61-100 with isHashKeys): 3 SimHash bit
    slices (16-bit key type), 4 weighted-MinHash counts (8-bit key type),
    5 one-bit weighted MinHash (16-bit), 7 cross-polytope.  wmh = (bounds,
    start, M) for 4 and 5."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, d = x.shape
    if keys_impl == 3:
        planes = normal_stream(tables * bits * d)
        return simhash(x, planes, tables, bits) & np.uint32(0xFFFF)
    if keys_impl in (4, 5):
        bounds, start, M = wmh
        if keys_impl == 4:
            return wmh_sketch(x, bounds, start, M, tables, 0, bit_stream(4096)) & np.uint32(0xFF)
        return wmh_sketch(x, bounds, start, M, tables, 1, bit_stream(4096), hbits=bits) & np.uint32(0xFFFF)
    if keys_impl == 7:
        P = pow2_at_least(d)
        return crosspoly(x, tables, sign_stream(tables * 3 * P)).astype(np.uint32)
    if keys_impl == 8:
        b = int(np.log2(2 * pow2_at_least(d)))
        assert bits % b == 0
        return crosspoly_dense(x, tables, bits // b)
    raise ValueError("keys implementation must be 3, 4, 5, 7 or 8")


def lsh_search(data, queries, k: int, implementation: int, keys_impl: int, sketch_dim: int, func: int,
               bits: int, tables: int, with_sketch: bool, C_: int = 4096):
    """(K,L) framework search (lshFramework.cuh:312-460) as sets: every
    query's candidates are the distinct points that share its key in at least
    one table; they are ranked by exact distance `func`, or (with_sketch and
    implementation >= 3) cut to the top C_ by sketch distance with ID ties and
    then ranked exactly.  Each hash family first preprocesses data and queries
    together, in the reference's order (keys, then sketches).
    Returns (ids, distances) [Q][calculate_k(k)]."""
    data = np.array(data, dtype=np.float32, copy=True, order="C")
    queries = np.array(queries, dtype=np.float32, copy=True, order="C")
    kq = calculate_k(k)

    def preprocess(family):
        nonlocal data, queries
        if family in (4, 5):
            data, queries = gj_transform(data, queries)
        elif family in (6, 7, 8):
            data, queries = normalize_rows(data), normalize_rows(queries)

    preprocess(keys_impl)
    wmh = wmh_bounds(data, queries) if keys_impl in (4, 5) else None
    dk = lsh_keys(data, keys_impl, tables, bits, wmh)
    qk = lsh_keys(queries, keys_impl, tables, bits, wmh)
    preprocess(implementation)
    sketched = with_sketch and implementation >= 3
    if sketched:
        if implementation == 3:
            planes = normal_stream(sketch_dim * 32 * data.shape[1])
            ds, qs = simhash(data, planes, sketch_dim), simhash(queries, planes, sketch_dim)
        elif implementation in (4, 5):
            b, st, M = wmh_bounds(data, queries)
            mode, H = implementation - 4, sketch_dim
            ds = wmh_sketch(data, b, st, M, H, mode, bit_stream(4096))
            qs = wmh_sketch(queries, b, st, M, H, mode, bit_stream(4096))
        else:
            planes = normal_stream(sketch_dim * data.shape[1], 1.0 / sketch_dim)
            ds, qs = jl(data, planes, sketch_dim), jl(queries, planes, sketch_dim)
    ids = np.empty((queries.shape[0], kq), dtype=np.int32)
    dist = np.empty((queries.shape[0], kq), dtype=np.float32)
    for q in range(queries.shape[0]):
        cand = lsh_candidates(dk, qk, q)
        if sketched and cand.size:
            c = min(C_, data.shape[0])
            if implementation == 6:
                sub = jl_topc(ds[cand], qs[q:q + 1], c)[0]
            else:
                sub = sketch_topc(1 if implementation == 4 else 0, ds[cand], qs[q:q + 1], c)[0][0]
            cand = cand[sub[sub >= 0]]
        row = np.full((1, max(cand.size, 1)), -1, dtype=np.int32)
        row[0, :cand.size] = cand
        i, dd = rerank(data, queries[q:q + 1], row, kq, func)
        ids[q], dist[q] = i[0], dd[0]
    return ids, dist


def goldstandard_port(data, queries, K: int, threads: int = 1):
    """Times goldstandard.cpp:136-173 (restated) over all given queries.
    Returns (seconds, ids[Q][K])."""
    data, dp = _f(data)
    queries, qp = _f(queries)
    ids = np.empty((queries.shape[0], K), dtype=np.int32)
    t = load().orc_goldstandard_port(dp, data.shape[0], qp, queries.shape[0], data.shape[1], int(K), int(threads),
                                     ids.ctypes.data_as(_I32))
    return float(t), ids


# ----------------------------------------------------------------- reference build

def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def _ref():
    lib = C.CDLL(REF_LIB)
    lib.ref_run_validation.argtypes = [C.c_char_p, _FP, C.c_void_p, C.c_int, C.c_int, C.c_int]
    lib.ref_random_vectors.argtypes = [_FP, C.c_int]
    lib.ref_random_vectors_sd.argtypes = [_FP, C.c_int, C.c_int]
    lib.ref_random_signs.argtypes = [_I32, C.c_int]
    return lib


def ref_random_vectors(n: int):
    out = np.empty(int(n), dtype=np.float32)
    _ref().ref_random_vectors(out.ctypes.data_as(_FP), int(n))
    return out


def ref_random_signs(n: int):
    out = np.empty(int(n), dtype=np.int32)
    _ref().ref_random_signs(out.ctypes.data_as(_I32), int(n))
    return out


def ref_run_validation(truth_file: str, points: np.ndarray, n_queries: int, k: int, report_k: int):
    """validation.cpp:201-208 on a truth file and a (Q, k) array of (ID, distance)."""
    pts = np.ascontiguousarray(points)
    container = np.zeros(2, dtype=np.float32)
    _ref().ref_run_validation(truth_file.encode(), container.ctypes.data_as(_FP), C.c_void_p(pts.ctypes.data),
                              int(n_queries), int(k), int(report_k))
    return float(container[0]), float(container[1])


def write_text_dataset(path: str, x: np.ndarray) -> None:
    """The `N\\nd\\n` then `id v1 ... vd ` per line format (DataParsing/Writer/writer.cpp:10-29)."""
    with open(path, "w") as f:
        f.write(f"{x.shape[0]}\n{x.shape[1]}\n")
        for i, row in enumerate(x):
            f.write(f"{i} " + " ".join(repr(float(v)) for v in row) + " \n")


def read_truth_file(path: str):
    """goldstandard.cpp:123-134: `n\\nk\\n` then per query `id:` and k lines `ID distance`."""
    with open(path) as f:
        tok = f.read().split()
    n, k = int(tok[0]), int(tok[1])
    ids = np.empty((n, k), dtype=np.int32)
    dist = np.empty((n, k), dtype=np.float32)
    p = 2
    for q in range(n):
        p += 1  # "q:"
        for j in range(k):
            ids[q, j] = int(tok[p])
            dist[q, j] = float(tok[p + 1])
            p += 2
    return ids, dist


def bucket_distribution(hashes):
    """statistics.cuh:10-16 bucketDistribution: res[hashes[i]] += 1 for every
    byte; all 256 values (the reference's result array has 255 entries)."""
    counts = [0] * 256
    for v in np.asarray(hashes, dtype=np.uint8).reshape(-1).tolist():
        counts[v] += 1
    return np.array(counts, dtype=np.int64)


def bucket_distribution_full_key(hashes, sketch_dim: int):
    """statisticsCpu.cpp:11-28 bucketDistributionFullKey: a map from each
    key's concatenated bitset<8> strings to its count, read in the map's
    string order.  Returns (keys [m][sketch_dim] uint8, counts [m])."""
    sd = int(sketch_dim)
    m = {}
    for row in np.asarray(hashes, dtype=np.uint8).reshape(-1, sd).tolist():
        key = "".join(format(b, "08b") for b in row)
        m[key] = m.get(key, 0) + 1
    order = sorted(m)
    keys = np.array([[int(k[8 * j:8 * j + 8], 2) for j in range(sd)] for k in order], dtype=np.uint8)
    return keys.reshape(len(order), sd), np.array([m[k] for k in order], dtype=np.int64)


def bucket_occupancy(keys, table_size: int, hist_len: int):
    """Bucket sizes of one table from its points' keys (the counts of the
    reference's bucket build, lshFramework.cuh:102-124): (buckets, non-empty
    buckets, points, largest bucket) and the histogram of sizes whose last
    bin collects the sizes >= hist_len - 1."""
    sizes = [0] * int(table_size)
    for k in np.asarray(keys).reshape(-1).tolist():
        sizes[k] += 1
    hist = [0] * max(int(hist_len), 0)
    for s in sizes:
        if hist:
            hist[min(s, len(hist) - 1)] += 1
    summary = {"buckets": len(sizes), "nonempty": sum(1 for s in sizes if s), "points": sum(sizes),
               "max_size": max(sizes) if sizes else 0}
    return summary, np.array(hist, dtype=np.int64)
