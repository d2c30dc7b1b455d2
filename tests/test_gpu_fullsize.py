"""Parity at the benchmark configurations' full sizes.

The oracle cannot sketch millions of rows in test time, so each test pins the
GPU's own sketches of every row on a random sample of rows (bit for bit
against the oracle), then checks, for a random sample of queries, that the
resident index returns exactly what the oracle's top-C and FP64 rerank give
over those sketches: the index's threshold pass (pivot cells or the sampled
full pass), selection and rerank are then checked at the real shapes.
"""
import numpy as np
import pytest

from lshongpu_amd import synth

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _close(a, b):
    np.testing.assert_allclose(a, b, rtol=RTOL, atol=1e-6)


@pytest.mark.timeout(600)
@pytest.mark.parametrize("cfg,words,C,nsample", [(1, 32, 1024, 64), (4, 16, 1024, 32)])
def test_simhash_index_full_size(lsh, orc, cfg, words, C, nsample):
    # cfg 1: 1.2M points, 10k queries, 1024-bit sketches (the benchmark);
    # cfg 4: 10M points, 2,048 queries, 512-bit sketches.
    data, queries = synth.config(cfg)
    N, d = data.shape
    Q = queries.shape[0]
    rng = np.random.default_rng(1000 + cfg)
    planes = orc.normal_stream(words * 32 * d)
    ds = lsh.simhash_sketch(data, words)
    rows = rng.choice(N, 256, replace=False)
    np.testing.assert_array_equal(ds[rows], orc.simhash(data[rows], planes, words))
    sample = np.sort(rng.choice(Q, nsample, replace=False))
    qs = orc.simhash(queries[sample], planes, words)
    np.testing.assert_array_equal(lsh.simhash_sketch(queries[sample], words), qs)
    with lsh.Index(data, implementation=3, sketchDim=words, distanceFunc=1, candidates=C) as ix:
        res = ix.search(queries, 32)
    cand, _ = orc.sketch_topc(0, ds, qs, C)
    ids, dist = orc.rerank(data, queries[sample], cand, 32, 1)
    np.testing.assert_array_equal(res["ID"][sample], ids)
    _close(res["distance"][sample], dist)


@pytest.mark.timeout(600)
def test_weighted_minhash_index_full_size(lsh, orc):
    # cfg 2: 1M sparse non-negative points, d = 128, 256 weighted-MinHash
    # values per point, top-32 by matching values, then the
    # generalized-Jaccard rerank (32 of its 5,000 queries).
    data, queries = synth.config(2)
    N, d = data.shape
    H, C, nsample = 256, 32, 32
    rng = np.random.default_rng(1002)
    # The data are non-negative, so the shift is zero whether or not the
    # queries take part; the index bounds the draws over its data alone.
    td, tq = orc.gj_transform(data, queries)
    ds = lsh.wmh_sketch(td, H)
    bounds, start, M = orc.wmh_bounds(td)
    bm = orc.bit_stream(4096)
    rows = rng.choice(N, 128, replace=False)
    np.testing.assert_array_equal(ds[rows], orc.wmh_sketch(td[rows], bounds, start, M, H, 0, bm))
    sample = np.sort(rng.choice(queries.shape[0], nsample, replace=False))
    qs = orc.wmh_sketch(tq[sample], bounds, start, M, H, 0, bm)
    # The match-count scan has no matrix-core form: the sampled queries are
    # searched (against all 1M points).
    with lsh.Index(data, implementation=4, sketchDim=H, distanceFunc=2, candidates=C) as ix:
        res = ix.search(queries[sample], 32)
    cand, _ = orc.sketch_topc(1, ds, qs, C)
    ids, dist = orc.rerank(td, tq[sample], cand, 32, 2)
    np.testing.assert_array_equal(res["ID"], ids)
    _close(res["distance"], dist)


@pytest.mark.timeout(600)
def test_lsh_framework_full_size(lsh, orc):
    # cfg 3: 4M clustered points, d = 128, (K,L) framework with one
    # cross-polytope key per table (2P = 256 buckets), L = 32 tables, exact
    # angular rescan of each query's bucket union (16 of its 50,000 queries).
    data, queries = synth.config(3)
    N, d = data.shape
    tables, bits, k, nsample = 32, 8, 32, 16
    rng = np.random.default_rng(1003)
    sample = np.sort(rng.choice(queries.shape[0], nsample, replace=False))
    qs = queries[sample]
    P = 128
    signs = orc.sign_stream(tables * 3 * P)
    xn, qn = orc.normalize_rows(data), orc.normalize_rows(qs)
    dk = lsh.crosspoly_keys(data, tables)
    rows = rng.choice(N, 256, replace=False)
    np.testing.assert_array_equal(dk[rows], orc.crosspoly(xn[rows], tables, signs))
    qk = orc.crosspoly(qn, tables, signs)
    np.testing.assert_array_equal(lsh.crosspoly_keys(qs, tables), qk)
    res = lsh.LSH(2, 7, k, d, nsample, N, data, qs, 1, 1, bits, tables, False)
    dk32, qk32 = dk.astype(np.uint32), qk.astype(np.uint32)
    for i in range(nsample):
        cand = orc.lsh_candidates(dk32, qk32, i)
        assert cand.size >= k
        ids, dist = orc.rerank(xn, qn[i:i + 1], cand[None, :], k, 1)
        np.testing.assert_array_equal(res.ids[i], ids[0])
        _close(res.distances[i], dist[0])


@pytest.mark.timeout(900)
def test_lsh_framework_dense_keys_full_size(lsh, orc):
    # cfg 3 as specified: 4M clustered points, d = 128, K = 2 dense-rotation
    # cross-polytope hashes per table (16 of the 32 key bits), L = 32 tables
    # sorted on the device, exact angular rescan (16 of the 50,000 queries).
    # Keys are pinned on 256 sampled rows and on the sampled queries; each
    # sampled query's result must be the oracle's FP64 top-k of its bucket
    # union.
    data, queries = synth.config(3)
    N, d = data.shape
    tables, K, bits, k, nsample = 32, 2, 16, 32, 16
    rng = np.random.default_rng(1004)
    sample = np.sort(rng.choice(queries.shape[0], nsample, replace=False))
    qs = queries[sample]
    G = orc.dense_rotations(tables, K, 128, d)
    xn, qn = orc.normalize_rows(data), orc.normalize_rows(qs)
    dk = lsh.crosspoly_dense_keys(data, tables, K)
    rows = rng.choice(N, 256, replace=False)
    np.testing.assert_array_equal(dk[rows], orc.crosspoly_dense(xn[rows], tables, K, G))
    qk = orc.crosspoly_dense(qn, tables, K, G)
    np.testing.assert_array_equal(lsh.crosspoly_dense_keys(qs, tables, K), qk)
    res = lsh.LSH(2, 8, k, d, nsample, N, data, qs, 1, 1, bits, tables, False)
    for i in range(nsample):
        cand = orc.lsh_candidates(dk, qk, i)
        assert cand.size >= k
        ids, dist = orc.rerank(xn, qn[i:i + 1], cand[None, :], k, 1)
        np.testing.assert_array_equal(res.ids[i], ids[0])
        _close(res.distances[i], dist[0])
