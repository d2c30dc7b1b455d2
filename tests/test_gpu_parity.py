"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Integer stages must be bit-exact (sketch bits, Hamming top-C IDs, keys);
returned IDs must equal the oracle's (FP64 distance, ID) order; returned
distances must be within 1e-5 relative of the FP64 oracle.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _close(a, b, rtol=RTOL):
    np.testing.assert_allclose(a, b, rtol=rtol, atol=1e-6)


def _unit(rng, n, d):
    x = rng.standard_normal((n, d)).astype(np.float32)
    return (x / np.linalg.norm(x, axis=1, keepdims=True)).astype(np.float32)


# --------------------------------------------------------------- exact scan

# The angular and -cos scans run on the FP32 matrix cores (d <= 128) unless
# LSHGPU_EXACT_VALU=1 is set when the index is built; both must give the
# oracle's (FP64, ID) top-k.
@pytest.fixture(params=["mfma", "valu"])
def exact_units(request, monkeypatch):
    if request.param == "valu":
        monkeypatch.setenv("LSHGPU_EXACT_VALU", "1")
    else:
        monkeypatch.delenv("LSHGPU_EXACT_VALU", raising=False)
    return request.param


@pytest.mark.parametrize("impl,func", [(2, 1), (2, 2), (2, 3), (1, 1)])
def test_exact_scan_matches_oracle(lsh, orc, impl, func, exact_units):
    rng = np.random.default_rng(100 + 10 * impl + func)
    N, Q, d, k = 5000, 70, 100, 32
    data = rng.standard_normal((N, d)).astype(np.float32)
    queries = rng.standard_normal((Q, d)).astype(np.float32)
    res = lsh.linearScans(impl, k, d, Q, N, data, queries, 0, func)
    if impl == 1:
        ids, dist = orc.exact_knn(data, queries, k, 0)
    elif func == 2:
        td, tq = orc.gj_transform(data, queries)
        ids, dist = orc.exact_knn(td, tq, k, 2)
    else:
        ids, dist = orc.exact_knn(data, queries, k, func)
    np.testing.assert_array_equal(res.ids, ids)
    _close(res.distances, dist)


def test_exact_scan_config0_size(lsh, orc, exact_units):
    from lshongpu_amd import synth

    data, queries = synth.config(0)
    res = lsh.linearScans(2, 32, 100, len(queries), len(data), data, queries, 0, 1)
    ids, dist = orc.exact_knn(data, queries, 32, 1)
    np.testing.assert_array_equal(res.ids, ids)
    _close(res.distances, dist)


def test_exact_scan_odd_dims_padding_and_k_rounding(lsh, orc, exact_units):
    rng = np.random.default_rng(7)
    N, Q, d = 20, 5, 37
    data = rng.standard_normal((N, d)).astype(np.float32)
    queries = rng.standard_normal((Q, d)).astype(np.float32)
    res = lsh.linearScans(2, 10, d, Q, N, data, queries, 0, 1)  # k=10 -> 32
    assert res.results.shape == (Q, 32)
    ids, dist = orc.exact_knn(data, queries, 32, 1)
    np.testing.assert_array_equal(res.ids, ids)
    assert (res.ids[:, N:] == -1).all()
    assert (res.distances[:, N:] == np.float32(2147483648.0)).all()
    _close(res.distances[:, :N], dist[:, :N])


def test_exact_scan_duplicates_break_ties_by_index(lsh, orc, exact_units):
    rng = np.random.default_rng(8)
    base = _unit(rng, 50, 16)
    data = np.concatenate([base[rng.integers(0, 50, 3000)], _unit(rng, 1000, 16)])
    queries = base[:9].copy()
    res = lsh.linearScans(2, 64, 16, 9, len(data), data, queries, 0, 1)
    ids, dist = orc.exact_knn(data, queries, 64, 1)
    np.testing.assert_array_equal(res.ids, ids)
    _close(res.distances, dist)


def test_exact_scan_large_k(lsh, orc, exact_units):
    rng = np.random.default_rng(9)
    data = _unit(rng, 6000, 24)
    queries = _unit(rng, 3, 24)
    res = lsh.linearScans(2, 1000, 24, 3, 6000, data, queries, 0, 3)
    ids, dist = orc.exact_knn(data, queries, 1024, 3)
    np.testing.assert_array_equal(res.ids, ids)
    _close(res.distances, dist)


def test_exact_scan_overflow_fallback_sorted_data(lsh, orc, exact_units):
    # Data ordered from farthest to nearest for every query: each scan round
    # appends its whole range, which overflows and takes the re-run path.
    rng = np.random.default_rng(10)
    q = _unit(rng, 1, 32)
    x = _unit(rng, 60000, 32)
    data = x[np.argsort(x @ q[0])].copy()  # cos ascending = distance descending
    queries = np.concatenate([q, q * 0.5 + _unit(rng, 1, 32) * 0.1]).astype(np.float32)
    res = lsh.linearScans(2, 32, 32, 2, len(data), data, queries, 0, 1)
    ids, dist = orc.exact_knn(data, queries, 32, 1)
    np.testing.assert_array_equal(res.ids, ids)
    _close(res.distances, dist)


def test_exact_scan_near_ties_escalate(lsh, orc, exact_units):
    # 200 points within 1e-7 of each other around the 32nd neighbour: the
    # FP32 margin cannot separate them, so the FP64 recheck has to decide.
    rng = np.random.default_rng(11)
    d = 64
    q = _unit(rng, 1, d)[0]
    far = _unit(rng, 5000, d)
    near = q + 0.05 * _unit(rng, 31, d)
    p = q + 0.2 * _unit(rng, 1, d)[0]
    tied = np.stack([p + 1e-7 * rng.standard_normal(d) for _ in range(200)])
    data = np.concatenate([far, near, tied]).astype(np.float32)
    data = data[rng.permutation(len(data))]
    res = lsh.linearScans(2, 64, d, 1, len(data), data, q[None].astype(np.float32), 0, 1)
    ids, dist = orc.exact_knn(data, q[None].astype(np.float32), 64, 1)
    np.testing.assert_array_equal(res.ids, ids)
    _close(res.distances, dist)


@pytest.mark.parametrize("d", [1, 3, 8, 9, 100, 126, 128, 129])
def test_exact_scan_matrix_cores_in_use_and_dims(lsh, orc, d, monkeypatch):
    # The matrix-core scan keeps the rows in operand order (tiles of 32 rows,
    # widths padded to 4 dimensions): its index holds that copy beside the rows, up to
    # d = 128; past it the vector scan runs.  Ragged N (not a multiple of 32)
    # and every padded width give the oracle's top-k.
    rng = np.random.default_rng(300 + d)
    N, Q, k = 1000 + 7, 45, 32
    data = rng.standard_normal((N, d)).astype(np.float32)
    queries = rng.standard_normal((Q, d)).astype(np.float32)
    sizes = {}
    for mode in ("mfma", "valu"):
        if mode == "valu":
            monkeypatch.setenv("LSHGPU_EXACT_VALU", "1")
        else:
            monkeypatch.delenv("LSHGPU_EXACT_VALU", raising=False)
        ix = lsh.Index(data, implementation=2, distanceFunc=1)
        sizes[mode] = ix.device_bytes()
        res = ix.search(queries, k)
        ids, dist = orc.exact_knn(data, queries, k, 1)
        np.testing.assert_array_equal(res["ID"], ids)
        _close(res["distance"], dist)
        ix.close()
    tiled = ((N + 31) // 32) * ((d + 3) // 4 * 2) * 64 * 4
    assert sizes["mfma"] - sizes["valu"] == (tiled if d <= 128 else 0)


# ------------------------------------------------------------ sketch stages

# Rows staged in LDS (d <= 255; full and partial words, word counts not a
# multiple of the 4 words a workgroup computes, rows not a multiple of its 64)
# and the unstaged kernel (d = 300).
@pytest.mark.parametrize("words,bits,n,d", [(32, 32, 3001, 100), (3, 32, 3001, 100), (12, 10, 3001, 100),
                                            (5, 32, 1027, 37), (4, 32, 130, 64), (1, 32, 2, 1),
                                            (7, 32, 515, 129), (2, 17, 70, 300)])
@pytest.mark.parametrize("valu", [False, True])
def test_simhash_bit_exact(lsh, orc, monkeypatch, words, bits, n, d, valu):
    # 32 bits a word and d <= 128 run on the matrix cores (bf16 splits, FP32
    # recheck near zero) unless LSHGPU_SIMHASH_VALU=1; the rest on the vector
    # units.  Both must give the oracle's FP32 bits.
    if valu:
        monkeypatch.setenv("LSHGPU_SIMHASH_VALU", "1")
    rng = np.random.default_rng(words * 100 + bits + d)
    x = rng.standard_normal((n, d)).astype(np.float32)
    planes = orc.normal_stream(words * bits * d)
    np.testing.assert_array_equal(lsh.simhash_sketch(x, words, bits), orc.simhash(x, planes, words, bits))


@pytest.mark.parametrize("d", [64, 100, 128])
def test_simhash_matrix_core_borderline_bits(lsh, orc, d):
    # Rows orthogonal (in FP64) to the first 20 hyperplanes: those projections
    # are rounding noise around zero, so every such bit takes the FP32
    # recheck; rows scaled from 2^-60 to 2^60 move the bound with the norm.
    rng = np.random.default_rng(d)
    words, n = 2, 700
    planes = orc.normal_stream(words * 32 * d).reshape(words * 32, d).astype(np.float64)
    x = rng.standard_normal((n, d))
    A = planes[:20]
    x = x - (x @ np.linalg.pinv(A)) @ A
    x[::7] *= 2.0 ** rng.integers(-60, 60, size=(len(x[::7]), 1))
    x = x.astype(np.float32)
    x[3] = 0.0
    want = orc.simhash(x, planes.astype(np.float32).ravel(), words)
    np.testing.assert_array_equal(lsh.simhash_sketch(x, words), want)


def test_jl_bit_exact(lsh, orc):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((1000, 50)).astype(np.float32)
    planes = orc.normal_stream(40 * 50, 1 / 40)
    np.testing.assert_array_equal(lsh.jl_sketch(x, 40), orc.jl(x, planes, 40))


def test_gj_transform_bit_exact(lsh, orc):
    rng = np.random.default_rng(4)
    data = (3 * rng.standard_normal((2000, 33))).astype(np.float32)
    queries = (3 * rng.standard_normal((17, 33))).astype(np.float32)
    gd, gq = lsh.gj_transform(data, queries)
    od, oq = orc.gj_transform(data, queries)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(gq, oq)


@pytest.mark.parametrize("d,tables", [(100, 8), (128, 3), (17, 5)])
def test_crosspoly_keys_bit_exact(lsh, orc, d, tables):
    rng = np.random.default_rng(d + tables)
    x = rng.standard_normal((2000, d)).astype(np.float32)
    P = 1 << int(np.ceil(np.log2(d)))
    signs = orc.sign_stream(tables * 3 * P)
    keys = lsh.crosspoly_keys(x, tables)
    np.testing.assert_array_equal(keys, orc.crosspoly(orc.normalize_rows(x), tables, signs))


@pytest.mark.parametrize("C", [1, 100, 1024])
def test_hamming_topc_bit_exact(lsh, orc, C):
    rng = np.random.default_rng(C)
    ds = (rng.integers(0, 2 ** 32, size=(20000, 5), dtype=np.uint64) & 0x0F0F0F0F).astype(np.uint32)
    qs = (rng.integers(0, 2 ** 32, size=(67, 5), dtype=np.uint64) & 0x0F0F0F0F).astype(np.uint32)
    ids, dist = lsh.hamming_topc(ds, qs, C)
    oids, odist = orc.sketch_topc(0, ds, qs, C)
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(dist, odist)


def test_hamming_topc_adversarial_order_overflows_and_recovers(lsh, orc):
    # Points ordered from the farthest to the nearest (in Hamming distance) of
    # query 0: every round finds better candidates than all before it, the
    # per-round capacity overflows, and the affected queries are re-run.
    rng = np.random.default_rng(77)
    ds = rng.integers(0, 2 ** 32, size=(60000, 8), dtype=np.uint64).astype(np.uint32)
    qs = rng.integers(0, 2 ** 32, size=(300, 8), dtype=np.uint64).astype(np.uint32)
    qs[1:40] = qs[0] ^ (1 << rng.integers(0, 32, size=(39, 8))).astype(np.uint32)
    h0 = np.unpackbits((ds ^ qs[0]).view(np.uint8), axis=1).sum(1)
    ds = ds[np.argsort(-h0, kind="stable")].copy()
    ids, dist = lsh.hamming_topc(ds, qs, 256)
    oids, odist = orc.sketch_topc(0, ds, qs, 256)
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(dist, odist)


@pytest.mark.gpu
def test_hamming_topc_sampled_threshold(lsh, orc):
    # Large enough for the sampled threshold (every 16th row): one matrix-core
    # pass appends the pairs at or below the sample's threshold distance.
    rng = np.random.default_rng(78)
    ds = rng.integers(0, 2 ** 32, size=(60000, 8), dtype=np.uint64).astype(np.uint32)
    qs = rng.integers(0, 2 ** 32, size=(100, 8), dtype=np.uint64).astype(np.uint32)
    qs[50:] = ds[rng.integers(0, 60000, 50)] ^ (1 << rng.integers(0, 32, size=(50, 8))).astype(np.uint32)
    ids, dist = lsh.hamming_topc(ds, qs, 64)
    oids, odist = orc.sketch_topc(0, ds, qs, 64)
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(dist, odist)


@pytest.mark.gpu
def test_hamming_topc_sampled_threshold_falls_back(lsh, orc):
    # Query 0: the sampled rows are its only near points, so the sample's
    # threshold admits fewer than C points.  Query 1: 600 unsampled copies of
    # it overflow the pass.  Both are redone with the rounds.
    rng = np.random.default_rng(79)
    ds = rng.integers(0, 2 ** 32, size=(60000, 8), dtype=np.uint64).astype(np.uint32)
    qs = rng.integers(0, 2 ** 32, size=(40, 8), dtype=np.uint64).astype(np.uint32)
    ds[np.arange(30) * 16] = qs[0] ^ np.array([1, 0, 0, 0, 0, 0, 0, 0], np.uint32)
    copies = np.arange(600) * 16 + 7
    ds[copies] = qs[1]
    ids, dist = lsh.hamming_topc(ds, qs, 64)
    oids, odist = orc.sketch_topc(0, ds, qs, 64)
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(dist, odist)
    assert (dist[1] == 0).all() and (ids[1] == copies[:64]).all()


@pytest.mark.gpu
def test_hamming_topc_two_sample_levels(lsh, orc):
    # 200k points, C = 16: samples of 12.5k and 782 rows, each level's
    # threshold coming from the one below.
    rng = np.random.default_rng(80)
    ds = rng.integers(0, 2 ** 32, size=(200000, 8), dtype=np.uint64).astype(np.uint32)
    qs = ds[rng.integers(0, 200000, 60)] ^ (1 << rng.integers(0, 32, size=(60, 8))).astype(np.uint32)
    ids, dist = lsh.hamming_topc(ds, qs, 16)
    oids, odist = orc.sketch_topc(0, ds, qs, 16)
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(dist, odist)


def _clustered_sketches(rng, n, words, centres, flip):
    """Rows near `centres` (each bit flipped with probability `flip`)."""
    c = centres[rng.integers(0, centres.shape[0], n)]
    bits = np.unpackbits(c.view(np.uint8), axis=1)
    bits ^= (rng.random(bits.shape) < flip).astype(np.uint8)
    return np.packbits(bits, axis=1).view(np.uint32).reshape(n, words)


@pytest.mark.gpu
@pytest.mark.parametrize("words,C", [(32, 256), (32, 64), (8, 100)])
def test_hamming_topc_pivot_cells(lsh, orc, words, C):
    # Clustered sketches, 150k points: the threshold pass runs over the pivot
    # cells and skips the tiles the triangle inequality rules out.  Some
    # queries are far from every cluster (wide thresholds), duplicates make
    # ties that cross cells.
    rng = np.random.default_rng(81 + words + C)
    centres = rng.integers(0, 2 ** 32, size=(300, words), dtype=np.uint64).astype(np.uint32)
    ds = _clustered_sketches(rng, 150000, words, centres, 0.08)
    dup = rng.integers(0, 150000, 3000)
    ds[rng.integers(0, 150000, 3000)] = ds[dup]
    qs = _clustered_sketches(rng, 300, words, centres, 0.08)
    qs[:20] = rng.integers(0, 2 ** 32, size=(20, words), dtype=np.uint64).astype(np.uint32)
    qs[20:30] = ds[dup[:10]]
    ids, dist = lsh.hamming_topc(ds, qs, C)
    oids, odist = orc.sketch_topc(0, ds, qs, C)
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(dist, odist)


@pytest.mark.gpu
def test_simhash_index_with_pivot_cells_matches_oracle(lsh, orc):
    # The resident index at 120k clustered points (pivot cells built) against
    # the oracle pipeline: exact Hamming top-C, then the exact rerank.
    rng = np.random.default_rng(82)
    N, Q, d, words, C, k = 120000, 200, 48, 16, 256, 32
    centres = rng.standard_normal((400, d)).astype(np.float32)
    data = centres[rng.integers(0, 400, N)] + 0.3 * rng.standard_normal((N, d)).astype(np.float32)
    queries = centres[rng.integers(0, 400, Q)] + 0.3 * rng.standard_normal((Q, d)).astype(np.float32)
    data /= np.linalg.norm(data, axis=1, keepdims=True)
    queries /= np.linalg.norm(queries, axis=1, keepdims=True)
    with lsh.Index(data, implementation=3, sketchDim=words, distanceFunc=1, candidates=C) as ix:
        res = ix.search(queries, k)
    planes = orc.normal_stream(words * 32 * d)
    cand, _ = orc.sketch_topc(0, orc.simhash(data, planes, words), orc.simhash(queries, planes, words), C)
    oids, odist = orc.rerank(data, queries, cand, k, 1)
    np.testing.assert_array_equal(res["ID"], oids)
    _close(res["distance"], odist)


@pytest.mark.gpu
@pytest.mark.parametrize("k", [1, 10, 100])
def test_simhash_index_cells_other_k(lsh, orc, k):
    # The pivot-cell index at the k values a serving load draws (1, 10, 100;
    # rounded to 32, 32, 128).  k = 100 runs the rerank as one kernel with the
    # FP64 recheck inside (the split form holds at most 64 results).
    rng = np.random.default_rng(83 + k)
    N, Q, d, words, C = 120000, 120, 48, 16, 256
    centres = rng.standard_normal((400, d)).astype(np.float32)
    data = centres[rng.integers(0, 400, N)] + 0.3 * rng.standard_normal((N, d)).astype(np.float32)
    queries = centres[rng.integers(0, 400, Q)] + 0.3 * rng.standard_normal((Q, d)).astype(np.float32)
    data /= np.linalg.norm(data, axis=1, keepdims=True)
    queries /= np.linalg.norm(queries, axis=1, keepdims=True)
    res = _index_vs_oracle(lsh, orc, data, queries, words, C, k)
    assert res.shape == (Q, lsh.calculateK(k))


@pytest.mark.gpu
@pytest.mark.parametrize("words", [48, 64])
def test_simhash_index_wide_sketches(lsh, orc, words):
    # 1536- and 2048-bit sketches (the reference's published rows reach 2048
    # bits).  Queries pointing away from the data put every candidate at a
    # Hamming distance above 1278, in the select histogram's last bin, which
    # is resolved by a radix select over the whole key.
    rng = np.random.default_rng(90 + words)
    N, Q, d, C, k = 30000, 24, 40, 300, 32
    centres = rng.standard_normal((50, d)).astype(np.float32)
    data = centres[rng.integers(0, 50, N)] + 0.3 * rng.standard_normal((N, d)).astype(np.float32)
    data /= np.linalg.norm(data, axis=1, keepdims=True)
    queries = _unit(rng, Q, d)
    queries[:12] = -data[rng.integers(0, N, 12)]
    res = _index_vs_oracle(lsh, orc, data, queries, words, C, k)
    assert res.shape == (Q, 32)


@pytest.mark.gpu
@pytest.mark.parametrize("words", [44, 48])
def test_hamming_topc_last_select_bin(lsh, orc, words):
    # Distances of 1279 and more share the select histogram's last bin.  The
    # data are one sketch c with about 3% of its bits flipped per row (so
    # thousands of rows tie at each distance), the first queries its
    # complement, so their whole top-C lies in that bin; random queries use
    # the middle bins.
    rng = np.random.default_rng(95 + words)
    N, C = 20000, 300
    p = 0.03
    c = rng.integers(0, 2 ** 32, size=words, dtype=np.uint64).astype(np.uint32)
    flips = np.packbits((rng.random((N, words * 32)) < p).astype(np.uint8), axis=1).view(np.uint32)
    ds = (c[None, :] ^ flips).astype(np.uint32)
    qs = rng.integers(0, 2 ** 32, size=(16, words), dtype=np.uint64).astype(np.uint32)
    qs[:10] = ~c
    qs[5:10, 0] ^= np.uint32(1) << np.arange(5, dtype=np.uint32)
    ids, dist = lsh.hamming_topc(ds, qs, C)
    oids, odist = orc.sketch_topc(0, ds, qs, C)
    assert (odist[:10] >= 1279).all()
    np.testing.assert_array_equal(ids, oids)
    np.testing.assert_array_equal(dist, odist)


def _index_vs_oracle(lsh, orc, data, queries, words, C, k):
    d = data.shape[1]
    with lsh.Index(data, implementation=3, sketchDim=words, distanceFunc=1, candidates=C) as ix:
        res = ix.search(queries, k)
    planes = orc.normal_stream(words * 32 * d)
    cand, _ = orc.sketch_topc(0, orc.simhash(data, planes, words), orc.simhash(queries, planes, words), C)
    oids, odist = orc.rerank(data, queries, cand, lsh.calculateK(k), 1)
    np.testing.assert_array_equal(res["ID"], oids)
    _close(res["distance"], odist)
    return res


@pytest.mark.gpu
def test_simhash_index_projection_modes_agree(lsh):
    # The index's data and query sketches on the matrix cores (projection 0)
    # and on the vector units (1) are the same bits, so searches agree.
    rng = np.random.default_rng(84)
    N, Q, d = 100000, 200, 100
    centres = rng.standard_normal((300, d)).astype(np.float32)
    data = centres[rng.integers(0, 300, N)] + 0.3 * rng.standard_normal((N, d)).astype(np.float32)
    queries = centres[rng.integers(0, 300, Q)] + 0.3 * rng.standard_normal((Q, d)).astype(np.float32)
    with lsh.Index(data, implementation=3, sketchDim=32, candidates=512, projection=0) as a, \
            lsh.Index(data, implementation=3, sketchDim=32, candidates=512, projection=1) as b:
        np.testing.assert_array_equal(a.search(queries, 32), b.search(queries, 32))


@pytest.mark.gpu
def test_simhash_index_redo_after_early_rerank(lsh, orc):
    # 1,500 copies of one point overflow its query's pass (more than 2C + 512
    # pairs at distance 0): the query is redone after its rerank was already
    # enqueued, and reranked again.
    rng = np.random.default_rng(83)
    N, Q, d, words, C, k = 120000, 640, 48, 16, 256, 32
    centres = rng.standard_normal((400, d)).astype(np.float32)
    data = centres[rng.integers(0, 400, N)] + 0.3 * rng.standard_normal((N, d)).astype(np.float32)
    data /= np.linalg.norm(data, axis=1, keepdims=True)
    copies = rng.choice(N, 1500, replace=False)
    data[copies] = data[copies[0]]
    queries = data[rng.integers(0, N, Q)] + 0.01 * rng.standard_normal((Q, d)).astype(np.float32)
    queries[0] = data[copies[0]]
    res = _index_vs_oracle(lsh, orc, data, queries, words, C, k)
    assert (res["ID"][0] == np.sort(copies)[:k]).all()


@pytest.mark.gpu
def test_simhash_index_uniform_data_falls_back(lsh, orc):
    # Unclustered data: the pivot cells cannot prune, the search falls back to
    # the sampled threshold and the full pass after the cells' rerank was
    # enqueued, and reranks again.
    rng = np.random.default_rng(84)
    data = _unit(rng, 100000, 32)
    queries = _unit(rng, 80, 32)
    _index_vs_oracle(lsh, orc, data, queries, 8, 256, 32)


@pytest.mark.gpu
def test_simhash_index_search_batches_through_graphs(lsh, orc):
    # Uneven batches (empty, 1, 2, odd sizes, one above the largest graph
    # class of 4096) through the captured graphs of the SimHash index with
    # pivot cells give every query what one ordinary search of all of them
    # gives, and the oracle's.  One query's pass overflows (1,500 copies of
    # its point at distance 0): the select flags it and it is redone after
    # the graphs.  A second call replays the graphs; device buffers in and
    # out; k = 10 shares the k = 32 graphs, k = 100 (rounded 128, the
    # rerank without the FP64 recheck kernel) has graphs of its own and
    # k = 200 runs ordinary searches.
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(86)
    N, d, words, C = 120000, 48, 16, 256
    sizes = [1, 0, 2, 37, 5, 4097, 300, 1, 64, 129]
    Q = sum(sizes)
    centres = rng.standard_normal((400, d)).astype(np.float32)
    data = centres[rng.integers(0, 400, N)] + 0.3 * rng.standard_normal((N, d)).astype(np.float32)
    data /= np.linalg.norm(data, axis=1, keepdims=True)
    copies = rng.choice(N, 1500, replace=False)
    data[copies] = data[copies[0]]
    queries = centres[rng.integers(0, 400, Q)] + 0.3 * rng.standard_normal((Q, d)).astype(np.float32)
    queries /= np.linalg.norm(queries, axis=1, keepdims=True)
    queries[5] = data[copies[0]]
    with lsh.Index(data, implementation=3, sketchDim=words, distanceFunc=1, candidates=C) as ix:
        want = ix.search(queries, 32)
        got = ix.search_batches(queries, sizes, 32)
        np.testing.assert_array_equal(got, want)
        # 8 batches of at most 4096 queries replayed graphs; the overflowing
        # query was redone
        st = ix.batch_stats()
        assert st["graph_batches"] == 8 and st["redone_queries"] >= 1, st
        np.testing.assert_array_equal(ix.search_batches(queries, sizes, 32), want)
        dq = torch.from_numpy(queries).cuda()
        dout = torch.empty((Q, 64), dtype=torch.int32, device="cuda")
        ix.search_batches(dq, sizes, 32, out=dout)
        np.testing.assert_array_equal(dout.cpu().numpy().view(lsh.POINT_DTYPE).reshape(Q, 32), want)
        np.testing.assert_array_equal(ix.search_batches(queries[:300], [150, 7, 143], 10),
                                      ix.search(queries[:300], 10))
        np.testing.assert_array_equal(ix.search_batches(queries[:300], [250, 50], 100),
                                      ix.search(queries[:300], 100))
        assert ix.batch_stats()["graph_batches"] == 2
        np.testing.assert_array_equal(ix.search_batches(queries[:300], [250, 50], 200),
                                      ix.search(queries[:300], 200))
        assert ix.batch_stats()["graph_batches"] == 0
    assert (want["ID"][5] == np.sort(copies)[:32]).all()
    planes = orc.normal_stream(words * 32 * d)
    cand, _ = orc.sketch_topc(0, orc.simhash(data, planes, words), orc.simhash(queries[:40], planes, words), C)
    oids, odist = orc.rerank(data, queries[:40], cand, 32, 1)
    np.testing.assert_array_equal(got["ID"][:40], oids)
    _close(got["distance"][:40], odist)


@pytest.mark.gpu
def test_simhash_index_search_batches_dense_cells_redone(lsh, orc):
    # Unclustered data: a captured batch finds the cells' second pass too
    # dense and is redone by an ordinary search (the sampled threshold and
    # the full pass); later calls on the index run ordinary searches.
    rng = np.random.default_rng(87)
    data = _unit(rng, 100000, 32)
    queries = _unit(rng, 300, 32)
    with lsh.Index(data, implementation=3, sketchDim=8, distanceFunc=1, candidates=256) as ix:
        want = ix.search(queries, 32)
        np.testing.assert_array_equal(ix.search_batches(queries, [100, 1, 199], 32), want)
        st = ix.batch_stats()
        assert st["graph_batches"] == 3 and st["redone_queries"] > 0, st
        np.testing.assert_array_equal(ix.search_batches(queries, [3, 297], 32), want)
        assert ix.batch_stats() == {"graph_batches": 0, "redone_queries": 0}
    planes = orc.normal_stream(8 * 32 * 32)
    cand, _ = orc.sketch_topc(0, orc.simhash(data, planes, 8), orc.simhash(queries, planes, 8), 256)
    oids, _ = orc.rerank(data, queries, cand, 32, 1)
    np.testing.assert_array_equal(want["ID"], oids)


def _rescaled_clusters(rng, N, Q, d, nc=400):
    """Clustered rows with norms spread over 1e-3..1e3 (the angular distance
    does not see the scale; the rerank's FP16 image stores unit rows)."""
    centres = rng.standard_normal((nc, d)).astype(np.float32)
    data = centres[rng.integers(0, nc, N)] + 0.3 * rng.standard_normal((N, d)).astype(np.float32)
    queries = centres[rng.integers(0, nc, Q)] + 0.3 * rng.standard_normal((Q, d)).astype(np.float32)
    data *= (10.0 ** rng.uniform(-3, 3, (N, 1))).astype(np.float32)
    queries *= (10.0 ** rng.uniform(-3, 3, (Q, 1))).astype(np.float32)
    return data, queries


@pytest.mark.gpu
@pytest.mark.parametrize("f32", ["0", "1"])
@pytest.mark.parametrize("d,C", [(37, 512), (64, 512), (40, 24)])
def test_simhash_index_rerank_filters_match_oracle(lsh, orc, monkeypatch, f32, d, C):
    # The cell index reranks with an FP16 image of the unit rows ("0") or the
    # FP32 rows ("1"); both recheck in FP64 every candidate within the error
    # band of the k-th and must give the oracle's result.  Odd d pads the
    # FP16 rows; near-duplicate groups (offsets of 1e-4) put hundreds of
    # candidates inside the FP16 band, exact copies make ties broken by ID.
    # With d % 4 == 0 the band's rows are staged through LDS in chunks (many
    # chunks at C = 512; two rows a chunk at C = 24, where fewer candidates
    # than k leave padded results).
    monkeypatch.setenv("LSHGPU_RERANK_F32", f32)
    rng = np.random.default_rng(85 + d + C)
    N, Q, words, k = 110000, 160, 16, 32
    data, queries = _rescaled_clusters(rng, N, Q, d)
    base = rng.integers(0, N, 8)
    for j, b in enumerate(base):
        grp = rng.choice(N, 300, replace=False)
        data[grp] = data[b] * (1 + 1e-4 * rng.standard_normal((300, d)).astype(np.float32))
        data[grp[:40]] = data[b]
        queries[j] = data[b] * 0.5
    _index_vs_oracle(lsh, orc, data, queries, words, C, k)


@pytest.mark.gpu
@pytest.mark.parametrize("d", [37, 100, 128, 130])
def test_simhash_index_fused_distances_match_oracle(lsh, orc, monkeypatch, d):
    # The cell scan packs each appended pair's approximate angular distance
    # (FP16 rows against the query's FP16 parts on the matrix cores, 16-bit
    # fixed point) into its key, and the filter reads it there instead of
    # gathering rows (d <= 128; d = 130 keeps the gather).  Both forms, and
    # k beyond the filter's split form (100), must give the oracle's result.
    monkeypatch.delenv("LSHGPU_RERANK_F32", raising=False)
    rng = np.random.default_rng(98 + d)
    N, Q, words, C = 110000, 160, 16, 512
    data, queries = _rescaled_clusters(rng, N, Q, d)
    for j, b in enumerate(rng.integers(0, N, 4)):
        grp = rng.choice(N, 200, replace=False)
        data[grp] = data[b] * (1 + 1e-4 * rng.standard_normal((200, d)).astype(np.float32))
        queries[j] = data[b] * 0.5
    # Norms whose FP32 squares underflow or overflow: the FP16 parts are
    # zero and the filter rechecks every candidate of these queries.
    queries[4] *= np.float32(1e-25)
    queries[5] *= np.float32(1e25)
    got = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("LSHGPU_FUSED", fused)
        got[fused] = _index_vs_oracle(lsh, orc, data, queries, words, C, 32)
    np.testing.assert_array_equal(got["1"], got["0"])
    monkeypatch.setenv("LSHGPU_FUSED", "1")
    _index_vs_oracle(lsh, orc, data, queries, words, C, 100)


@pytest.mark.gpu
def test_simhash_index_rerank_64bit_image_offsets(lsh, orc, monkeypatch):
    # The FP16 filter addresses rows by 32-bit byte offsets below 4 GiB of
    # image and by 64-bit piece indices above; LSHGPU_RERANK_OFF64=1 takes
    # the second path on a small index, which must give the same result.
    monkeypatch.delenv("LSHGPU_RERANK_F32", raising=False)
    monkeypatch.setenv("LSHGPU_RERANK_OFF64", "1")
    rng = np.random.default_rng(97)
    N, Q, d, words, C, k = 110000, 160, 100, 16, 512, 32
    data, queries = _rescaled_clusters(rng, N, Q, d)
    _index_vs_oracle(lsh, orc, data, queries, words, C, k)


@pytest.mark.gpu
def test_simhash_index_zero_row_keeps_fp32_filter(lsh, orc, monkeypatch):
    # A zero row has no unit direction: the index keeps the FP32 filter (no
    # FP16 image, so fewer device bytes) and still matches the oracle.
    monkeypatch.delenv("LSHGPU_RERANK_F32", raising=False)
    rng = np.random.default_rng(86)
    N, Q, d, words, C, k = 100000, 64, 40, 16, 256, 32
    data, queries = _rescaled_clusters(rng, N, Q, d)
    with lsh.Index(data, implementation=3, sketchDim=words, distanceFunc=1, candidates=C) as ix:
        with_half = ix.device_bytes()
    data[7] = 0.0
    with lsh.Index(data, implementation=3, sketchDim=words, distanceFunc=1, candidates=C) as ix:
        assert with_half - ix.device_bytes() >= 0.9 * N * 40 * 2
    _index_vs_oracle(lsh, orc, data, queries, words, C, k)


# ------------------------------------------------------------ full pipelines

def test_simhash_pipeline_matches_oracle(lsh, orc):
    rng = np.random.default_rng(21)
    N, Q, d, words, C, k = 20000, 40, 64, 8, 256, 32
    data = _unit(rng, N, d)
    queries = _unit(rng, Q, d)
    res = lsh.linearScans(3, k, d, Q, N, data, queries, words, 1, candidates=C)
    planes = orc.normal_stream(words * 32 * d)
    cand, _ = orc.sketch_topc(0, orc.simhash(data, planes, words), orc.simhash(queries, planes, words), C)
    ids, dist = orc.rerank(data, queries, cand, k, 1)
    np.testing.assert_array_equal(res.ids, ids)
    _close(res.distances, dist)


@pytest.mark.parametrize("impl", [3, 4, 5])
@pytest.mark.parametrize("C", [None, 3000])
def test_sketch_scans_default_candidates(lsh, orc, impl, C):
    # The drop-in call without options uses C = 4096, the reference's rerank
    # queue width; candidate sets above 2048 (4096 and 3000) take the sorted
    # float-key select instead of the integer-key one.
    rng = np.random.default_rng(60 + impl + (C or 0))
    N, Q, d, k = 20000, 12, 64, 32
    Ceff = C or 4096
    kw = {} if C is None else {"candidates": C}
    if impl == 3:
        sk = 8
        data, queries = _unit(rng, N, d), _unit(rng, Q, d)
        res = lsh.linearScans(3, k, d, Q, N, data, queries, sk, 1, **kw)
        planes = orc.normal_stream(sk * 32 * d)
        ds, qs = orc.simhash(data, planes, sk), orc.simhash(queries, planes, sk)
        td, tq, func, kind = data, queries, 1, 0
    else:
        sk = 24 if impl == 4 else 8
        data, queries = _wmh_rows(rng, N, d, 0.3), _wmh_rows(rng, Q, d, 0.3)
        res = lsh.linearScans(impl, k, d, Q, N, data, queries, sk, 2, **kw)
        td, tq = orc.gj_transform(data, queries)
        bounds, start, M = orc.wmh_bounds(td, tq)
        bm = orc.bit_stream(4096)
        ds = orc.wmh_sketch(td, bounds, start, M, sk, impl - 4, bm)
        qs = orc.wmh_sketch(tq, bounds, start, M, sk, impl - 4, bm)
        func, kind = 2, 1 if impl == 4 else 0
    cand, _ = orc.sketch_topc(kind, ds, qs, Ceff)
    ids, dist = orc.rerank(td, tq, cand, k, func)
    np.testing.assert_array_equal(res.ids, ids)
    _close(res.distances, dist)


def test_jl_pipeline_matches_oracle(lsh, orc):
    rng = np.random.default_rng(22)
    N, Q, d, sdim, C, k = 8000, 20, 48, 24, 128, 32
    data = rng.standard_normal((N, d)).astype(np.float32)
    queries = rng.standard_normal((Q, d)).astype(np.float32)
    res = lsh.linearScans(6, k, d, Q, N, data, queries, sdim, 1, candidates=C)
    nd, nq = orc.normalize_rows(data), orc.normalize_rows(queries)
    planes = orc.normal_stream(sdim * d, 1 / sdim)
    cand = orc.jl_topc(orc.jl(nd, planes, sdim), orc.jl(nq, planes, sdim), C)
    ids, dist = orc.rerank(nd, nq, cand, k, 1)
    np.testing.assert_array_equal(res.ids, ids)
    _close(res.distances, dist)


def test_jl_pipeline_near_ties_match_oracle(lsh, orc):
    # Rows repeated with last-bit perturbations: their JL sketch distances
    # differ by a few FP32 roundings, so the top-C set depends on the sums
    # being fl(t*t) plus a separate add (distanceFunctions.cuh:87-99); a fused
    # sum changes the set for most of these queries (checked on the CPU).
    # k = C makes the returned IDs the whole candidate set.
    rng = np.random.default_rng(41)
    d, sdim, C = 48, 24, 64
    base = rng.standard_normal((30, d)).astype(np.float32)
    data = np.repeat(base, 100, axis=0)
    data = (data * (1 + rng.integers(-3, 4, size=data.shape) * 2.0 ** -23)).astype(np.float32)
    queries = np.concatenate([base[:3] * 1.0001, rng.standard_normal((5, d))]).astype(np.float32)
    N, Q = len(data), len(queries)
    res = lsh.linearScans(6, C, d, Q, N, data, queries, sdim, 1, candidates=C)
    nd, nq = orc.normalize_rows(data), orc.normalize_rows(queries)
    planes = orc.normal_stream(sdim * d, 1 / sdim)
    cand = orc.jl_topc(orc.jl(nd, planes, sdim), orc.jl(nq, planes, sdim), C)
    ids, dist = orc.rerank(nd, nq, cand, C, 1)
    np.testing.assert_array_equal(res.ids, ids)
    _close(res.distances, dist)


def test_index_search_repeatable_and_batch_independent(lsh):
    rng = np.random.default_rng(23)
    data = _unit(rng, 30000, 100)
    queries = _unit(rng, 300, 100)
    with lsh.Index(data, implementation=3, sketchDim=16, candidates=512) as ix:
        a = ix.search(queries, 32)
        b = ix.search(queries, 32)
        c = ix.search(queries[:7], 32)
        one = ix.search(queries[5:6], 32)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a[:7], c)
    np.testing.assert_array_equal(a[5:6], one)


def test_index_device_pointers(lsh):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(24)
    data = _unit(rng, 10000, 64)
    queries = _unit(rng, 50, 64)
    host = lsh.Index(data, implementation=2, distanceFunc=1).search(queries, 32)
    dd = torch.from_numpy(data).cuda()
    dq = torch.from_numpy(queries).cuda()
    out = torch.empty((50, 32 * 2), dtype=torch.int32, device="cuda")
    with lsh.Index(dd, implementation=2, distanceFunc=1) as ix:
        ix.search(dq, 32, out=out)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(lsh.POINT_DTYPE).reshape(50, 32)
    np.testing.assert_array_equal(got, host)


@pytest.mark.parametrize("fused", ["1", "0"])
def test_index_searches_on_streams_are_ordered(lsh, monkeypatch, fused):
    # Device results on a caller's stream: the search may return before its
    # rerank has run.  Back-to-back searches on two streams into two outputs
    # are ordered by the index, and each matches the synchronous search; the
    # FP64 recheck runs as its own kernel on the pivot-cell index, after the
    # split filter (fused = "0", profiling mode 2 times the filter only) or
    # after the select that wrote the band from fused keys ("1", no filter
    # kernel; mode 3 times the cell search's appending pass only).
    monkeypatch.setenv("LSHGPU_FUSED", fused)
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(25)
    N, Q, d = 120000, 300, 48
    centres = rng.standard_normal((400, d)).astype(np.float32)
    data = centres[rng.integers(0, 400, N)] + 0.3 * rng.standard_normal((N, d)).astype(np.float32)
    queries = centres[rng.integers(0, 400, Q)] + 0.3 * rng.standard_normal((Q, d)).astype(np.float32)
    dd, dq = torch.from_numpy(data).cuda(), torch.from_numpy(queries).cuda()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    with lsh.Index(dd, implementation=3, sketchDim=16, distanceFunc=1, candidates=256) as ix:
        want = ix.search(queries, 32)
        ix.set_profiling(2 if fused == "0" else 3)
        outs = []
        for rep in range(3):
            o1 = torch.full((Q, 64), -7, dtype=torch.int32, device="cuda")
            o2 = torch.full((Q, 64), -7, dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()  # the fills ran on torch's current stream
            ix.search(dq, 32, out=o1, stream=s1.cuda_stream)
            ix.search(dq, 32, out=o2, stream=s2.cuda_stream)
            outs += [(o1, s1), (o2, s2)]
        pt = ix.phase_times()
        ix.set_profiling(False)
    for o, st in outs:
        st.synchronize()
        got = o.cpu().numpy().view(lsh.POINT_DTYPE).reshape(Q, 32)
        np.testing.assert_array_equal(got, want)
    assert pt["scan_ms"] == 0 and pt["recheck_ms"] == 0
    if fused == "0":
        assert pt["final_ms"] > 0 and pt["pass2_ms"] == 0
    else:
        assert pt["final_ms"] == 0 and pt["pass2_ms"] > 0 and pt["pass2_tiles"] > 0


@pytest.mark.parametrize("kw", [dict(implementation=1), dict(implementation=2, distanceFunc=2),
                                dict(implementation=3, sketchDim=16), dict(implementation=4, sketchDim=64),
                                dict(implementation=6, sketchDim=32),
                                dict(framework=1, implementation=2, keysImplementation=3, bucketKeyBits=6,
                                     tables=4)])
def test_index_device_buffers_match_host(lsh, kw):
    # Device queries are read in place when no transform applies (and copied
    # otherwise), device results are written in place: both give what host
    # buffers give, the queries are left unchanged, and a smaller batch
    # between two full ones (the query-sketch buffer keeps zero padding rows)
    # changes nothing.
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(31)
    N, Q, d = 5000, 96, 24
    data = np.abs(rng.standard_normal((N, d))).astype(np.float32)
    queries = np.abs(rng.standard_normal((Q, d))).astype(np.float32)
    dq = torch.from_numpy(queries).cuda()
    s = torch.cuda.Stream()
    with lsh.Index(data, **kw) as ix:
        for nq in (Q, 37, Q):
            want = ix.search(queries[:nq], 32)
            out = torch.full((nq, 64), -7, dtype=torch.int32, device="cuda")
            torch.cuda.synchronize()
            ix.search(dq[:nq], 32, out=out, stream=s.cuda_stream)
            s.synchronize()
            np.testing.assert_array_equal(out.cpu().numpy().view(lsh.POINT_DTYPE).reshape(nq, 32), want)
    np.testing.assert_array_equal(dq.cpu().numpy(), queries)


# ------------------------------------------------------------ weighted MinHash

def _wmh_rows(rng, n, d, density):
    x = rng.exponential(1.0, (n, d)).astype(np.float32)
    x *= rng.random((n, d)) < density
    x[np.arange(n), rng.integers(0, d, n)] += 0.5  # no empty rows
    return x


@pytest.mark.gpu
@pytest.mark.parametrize("one_bit", [False, True])
@pytest.mark.parametrize("d,density", [(40, 0.6), (300, 0.01)])
def test_wmh_sketch_bit_exact(lsh, orc, one_bit, d, density):
    """Sparse rows make most counts exceed the precomputed draw table."""
    rng = np.random.default_rng(31 + d)
    raw = _wmh_rows(rng, 700, d, density)
    x, _ = orc.gj_transform(raw, raw[:0])
    sk = 3 if one_bit else 40
    got = lsh.wmh_sketch(x, sk, one_bit)
    bounds, start, M = orc.wmh_bounds(x)
    want = orc.wmh_sketch(x, bounds, start, M, sk, int(one_bit), orc.bit_stream(4096))
    np.testing.assert_array_equal(got, want)
    if not one_bit and density < 0.1:
        assert (want > 256).mean() > 0.2


@pytest.mark.gpu
@pytest.mark.parametrize("impl", [4, 5])
def test_wmh_pipeline_matches_oracle(lsh, orc, impl):
    rng = np.random.default_rng(40 + impl)
    N, Q, d, C, k = 6000, 24, 64, 200, 32
    sk = 24 if impl == 4 else 8
    data = _wmh_rows(rng, N, d, 0.3)
    queries = _wmh_rows(rng, Q, d, 0.3)
    res = lsh.linearScans(impl, k, d, Q, N, data, queries, sk, 2, candidates=C)
    td, tq = orc.gj_transform(data, queries)
    bounds, start, M = orc.wmh_bounds(td, tq)
    bm = orc.bit_stream(4096)
    mode = impl - 4
    ds = orc.wmh_sketch(td, bounds, start, M, sk, mode, bm)
    qs = orc.wmh_sketch(tq, bounds, start, M, sk, mode, bm)
    cand, _ = orc.sketch_topc(1 if impl == 4 else 0, ds, qs, C)
    ids, dist = orc.rerank(td, tq, cand, k, 2)
    np.testing.assert_array_equal(res.ids, ids)
    _close(res.distances, dist)


# ------------------------------------------- dense-rotation cross-polytope

@pytest.mark.gpu
@pytest.mark.parametrize("d,tables,K", [(128, 4, 2), (100, 3, 2), (128, 2, 4), (17, 2, 3), (64, 3, 1), (5, 2, 2)])
def test_crosspoly_dense_keys_bit_exact(lsh, orc, d, tables, K):
    # Matrix-core products with the FP32 recheck of every ambiguous hash
    # against the oracle's FP32 definition, bit for bit; duplicates, a zero
    # row (NaN after normalisation: every coordinate rechecked) and rows that
    # differ in the last bit of one coordinate.
    rng = np.random.default_rng(d * 10 + tables + K)
    n = 5003
    x = rng.standard_normal((n, d)).astype(np.float32)
    x[10:20] = x[0]
    x[20] = 0.0
    x[21:40] = x[1]
    x[21:40, 0] = np.nextafter(x[21:40, 0], np.float32(np.inf))
    keys = lsh.crosspoly_dense_keys(x, tables, K)
    want = orc.crosspoly_dense(orc.normalize_rows(x), tables, K)
    np.testing.assert_array_equal(keys, want)


@pytest.mark.gpu
@pytest.mark.parametrize("N,L,bits", [(5000, 3, 8), (70001, 2, 16), (4099, 4, 24), (33000, 2, 32), (1, 1, 5)])
def test_sort_table_keys_stable(lsh, N, L, bits):
    rng = np.random.default_rng(N + bits)
    hi = min(2 ** bits, 2 ** 32)
    keys = rng.integers(0, max(2, hi // 64 if bits > 8 else hi), size=(L, N), dtype=np.uint64).astype(np.uint32)
    keys[:, ::7] = keys[:, :1]
    if bits < 32:
        keys[0, ::5] |= np.uint32(1) << np.uint32(bits)  # bits above the sorted ones keep their order
    sk, ids = lsh.sort_table_keys(keys, bits)
    mask = np.uint32((1 << bits) - 1) if bits < 32 else np.uint32(0xFFFFFFFF)
    for t in range(L):
        order = np.argsort(keys[t] & mask, kind="stable")
        np.testing.assert_array_equal(ids[t], order.astype(np.uint32))
        np.testing.assert_array_equal(sk[t], keys[t][order])


# ------------------------------------------------------------ (K,L) framework

LSH_CASES = [
    # implementation, keys, runWithSketchedData, distanceFunc, bucketKeyBits, tables, k, data kind
    (2, 3, False, 1, 10, 8, 32, "unit"),
    (2, 3, False, 3, 4, 3, 100, "unit"),        # 16 buckets per table: large candidate sets
    (2, 7, False, 1, 8, 6, 32, "unit"),
    (2, 5, False, 2, 6, 5, 32, "wmh"),
    (2, 4, False, 2, 8, 4, 32, "wmh"),
    (3, 3, True, 1, 12, 10, 32, "unit"),
    (4, 4, True, 2, 8, 6, 32, "wmh"),           # both families shift and normalise
    (5, 7, True, 2, 8, 4, 32, "wmh"),
    (2, 8, False, 1, 14, 6, 32, "unit"),        # dense-rotation keys, K = 2 (P = 64, 7-bit fields)
    (3, 8, True, 1, 14, 8, 32, "unit"),
    (2, 8, False, 3, 7, 10, 100, "unit"),       # K = 1: 128 buckets per table, large candidate sets
    (2, 3, False, 1, 20, 5, 32, "unit"),        # wider than 16 bits: sorted tables for a reference family
    (2, 7, False, 1, 18, 4, 32, "unit"),
    (6, 3, True, 3, 10, 6, 32, "unit"),
]


@pytest.mark.gpu
@pytest.mark.parametrize("impl,keys,sk,func,bits,tables,k,kind", LSH_CASES)
def test_lsh_framework_matches_oracle(lsh, orc, impl, keys, sk, func, bits, tables, k, kind):
    rng = np.random.default_rng(50 + impl * 7 + keys + bits)
    N, Q, d = 5000, 24, 64
    if kind == "unit":
        data, queries = _unit(rng, N, d), _unit(rng, Q, d)
    else:
        data, queries = _wmh_rows(rng, N, d, 0.3), _wmh_rows(rng, Q, d, 0.3)
    sketch_dim = {3: 4, 4: 16, 5: 2, 6: 16}.get(impl, 1)
    C = 300
    res = lsh.LSH(impl, keys, k, d, Q, N, data, queries, sketch_dim, func, bits, tables, sk, candidates=C)
    ids, dist = orc.lsh_search(data, queries, k, impl, keys, sketch_dim, func, bits, tables, sk, C)
    np.testing.assert_array_equal(res.ids, ids)
    _close(res.distances, dist)


@pytest.mark.gpu
def test_lsh_index_matches_one_shot_and_is_batch_independent(lsh, orc):
    rng = np.random.default_rng(61)
    N, Q, d = 20000, 64, 48
    data, queries = _unit(rng, N, d), _unit(rng, Q, d)
    cfg = dict(implementation=2, distanceFunc=1, framework=1, keysImplementation=3, bucketKeyBits=11, tables=12)
    with lsh.Index(data, **cfg) as ix:
        a = ix.search(queries, 32)
        b = ix.search(queries[10:11], 32)
    np.testing.assert_array_equal(a[10:11], b)
    # Without preprocessing the index and the one-shot call agree exactly.
    one = lsh.LSH(2, 3, 32, d, Q, N, data, queries, 1, 1, 11, 12, False)
    np.testing.assert_array_equal(a["ID"], one.ids)
    ids, _ = orc.lsh_search(data, queries, 32, 2, 3, 1, 1, 11, 12, False)
    np.testing.assert_array_equal(a["ID"], ids)


@pytest.mark.gpu
def test_lsh_rejects_keys_wider_than_the_table(lsh):
    rng = np.random.default_rng(62)
    data, queries = _unit(rng, 100, 64), _unit(rng, 4, 64)
    with pytest.raises(lsh.LshError):
        lsh.LSH(2, 7, 32, 64, 4, 100, data, queries, 1, 1, 6, 4, False)  # 2P = 128 > 2^6
    with pytest.raises(lsh.LshError):
        lsh.LSH(2, 4, 32, 64, 4, 100, data, queries, 1, 1, 6, 4, False)  # 8-bit counts > 2^6


@pytest.mark.gpu
def test_autotune_picks_fastest_config_meeting_recall(lsh):
    # Three (K,L) settings of the dense-rotation keys: the autotuner's recall
    # must equal the recall the chosen index gives, its choice must reach the
    # target and be the fastest that does or, among those within 3% of the
    # fastest, the most accurate; a second call is served from the cache with
    # the same answer.
    rng = np.random.default_rng(77)
    N, d, k = 40000, 64, 32
    centres = rng.standard_normal((100, d)).astype(np.float32)
    data = centres[rng.integers(0, 100, N)] + 0.25 * rng.standard_normal((N, d)).astype(np.float32)
    queries = centres[rng.integers(0, 100, 64)] + 0.25 * rng.standard_normal((64, d)).astype(np.float32)
    base = dict(implementation=2, distanceFunc=1, framework=1, keysImplementation=8)
    cands = [dict(base, bucketKeyBits=7, tables=8), dict(base, bucketKeyBits=14, tables=8),
             dict(base, bucketKeyBits=28, tables=2)]
    best, reps = lsh.autotune(data, queries, k, cands, target_recall=0.8)
    assert all(r["status"] == 0 for r in reps)
    truth = lsh.linearScans(2, k, d, len(queries), N, data, queries, 0, 1).ids
    for c, r in zip(cands, reps):
        with lsh.Index(data, **c) as ix:
            got = ix.search(queries, k)["ID"]
        rec = np.mean([len(set(a) & set(b)) / k for a, b in zip(got, truth)])
        assert abs(rec - r["recall"]) < 1e-12
    ok = [i for i, r in enumerate(reps) if r["recall"] >= 0.8]
    fastest = min(reps[i]["search_ms"] for i in ok)
    near = [i for i in ok if reps[i]["search_ms"] <= fastest * 1.03]
    assert best in near and reps[best]["recall"] == max(reps[i]["recall"] for i in near)
    again, reps2 = lsh.autotune(data, queries, k, cands, target_recall=0.8)
    assert again == best and reps2 == reps


def _dense_lsh_data(rng, N, Q, d):
    centres = rng.standard_normal((80, d)).astype(np.float32)
    data = centres[rng.integers(0, 80, N)] + 0.3 * rng.standard_normal((N, d)).astype(np.float32)
    queries = centres[rng.integers(0, 80, Q)] + 0.3 * rng.standard_normal((Q, d)).astype(np.float32)
    return data, queries


@pytest.mark.gpu
@pytest.mark.parametrize("keys,bits", [(8, 14), (3, 12), (7, 8)])
def test_search_batches_equal_one_search_per_batch(lsh, keys, bits):
    # Uneven batches (empty, 1, 2, ..., and one above the largest graph
    # class) through the captured graphs (keys 8 and 3; keys 7 runs them as
    # ordinary searches) give every query what one search of all of them
    # gives; a second call replays the graphs; device buffers in and out.
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(500 + keys)
    N, d = 30000, 64
    sizes = [1, 0, 2, 37, 5, 4097, 300, 1, 64, 129]
    data, queries = _dense_lsh_data(rng, N, sum(sizes), d)
    with lsh.Index(data, implementation=2, distanceFunc=1, framework=1, keysImplementation=keys,
                   bucketKeyBits=bits, tables=8) as ix:
        want = ix.search(queries, 32)
        got = ix.search_batches(queries, sizes, 32)
        np.testing.assert_array_equal(got, want)
        again = ix.search_batches(queries, sizes, 32)
        np.testing.assert_array_equal(again, want)
        dq = torch.from_numpy(queries).cuda()
        dout = torch.empty((len(queries), 64), dtype=torch.int32, device="cuda")
        ix.search_batches(dq, sizes, 32, out=dout)
        np.testing.assert_array_equal(dout.cpu().numpy().view(lsh.POINT_DTYPE).reshape(len(queries), 32), want)
        # k = 100 (rounded to 128) captures graphs of its own
        np.testing.assert_array_equal(ix.search_batches(queries[:500], [250, 250], 100), ix.search(queries[:500], 100))


@pytest.mark.gpu
def test_search_batches_rerun_of_capped_work_list(lsh, orc, monkeypatch):
    # A work-list capacity of 3 items leaves most queries incomplete: they
    # are flagged and re-run, and the results still equal the oracle's.
    rng = np.random.default_rng(510)
    N, d = 20000, 64
    data, queries = _dense_lsh_data(rng, N, 40, d)
    monkeypatch.setenv("LSHGPU_WORKLIST_CAP", "3")
    res = lsh.LSH(2, 8, 32, d, len(queries), N, data, queries, 1, 1, 14, 6, False)
    with lsh.Index(data, implementation=2, distanceFunc=1, framework=1, keysImplementation=8, bucketKeyBits=14,
                   tables=6) as ix:
        got = ix.search_batches(queries, [7, 33], 32)
    ids, dist = orc.lsh_search(data, queries, 32, 2, 8, 1, 1, 14, 6, False)
    np.testing.assert_array_equal(res.ids, ids)
    np.testing.assert_array_equal(got["ID"], ids)
    _close(got["distance"], dist)
