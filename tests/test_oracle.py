"""Pins of the CPU oracle (runs without a GPU).

Each oracle function is checked against something that is not itself: the
reference's own code compiled into oracle/_ref/ and the fixtures it wrote
(tests/golden/), an independent restatement of a published algorithm, or
brute force / closed forms in numpy.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN


# ----------------------------------------------------------------- calculateK

@pytest.mark.parametrize("k,expect", [(1, 32), (31, 32), (32, 32), (33, 64), (64, 64), (100, 128), (1024, 1024),
                                      (1025, 1056)])
def test_calculate_k_closed_form(orc, k, expect):
    # launchHelper.cuh:21-25: smallest multiple of 32 that is >= k
    assert orc.calculate_k(k) == expect == int(math.ceil(k / 32) * 32)


# ------------------------------------------------------------- random streams
# Independent restatement of libstdc++'s minstd_rand0 + generate_canonical +
# Marsaglia polar normal_distribution and the downscaling
# uniform_int_distribution, in plain Python.

class _MinStd0:
    def __init__(self):
        self.x = 1

    def __call__(self):
        self.x = (16807 * self.x) % 2147483647
        return self.x


def _canonical(g):
    r = 2147483646.0
    s = float(g() - 1)
    s = s + float(g() - 1) * r
    t = r * r
    v = s / t
    return v if v < 1.0 else math.nextafter(1.0, 0.0)


def _normals(n, stddev=1.0):
    g, out, saved = _MinStd0(), [], None
    while len(out) < n:
        if saved is not None:
            v, saved = saved, None
        else:
            while True:
                x = 2.0 * _canonical(g) - 1.0
                y = 2.0 * _canonical(g) - 1.0
                r2 = x * x + y * y
                if not (r2 > 1.0 or r2 == 0.0):
                    break
            mult = math.sqrt(-2 * math.log(r2) / r2)
            saved = x * mult
            v = y * mult
        out.append(np.float32(v * stddev + 0.0))
    return np.array(out, dtype=np.float32)


def _uniform_int(n, hi):
    g, out = _MinStd0(), []
    uerange = hi + 1
    scaling = 2147483645 // uerange
    past = uerange * scaling
    while len(out) < n:
        u = g() - 1
        if u >= past:
            continue
        out.append(u // scaling)
    return np.array(out)


def test_canonical_tmp_is_the_long_double_product():
    # __tmp *= __r is computed in long double and rounded to double once
    assert 2147483646.0 * 2147483646.0 == float(2147483646 ** 2)


def test_normal_stream_restated(orc):
    np.testing.assert_array_equal(orc.normal_stream(2000), _normals(2000))
    np.testing.assert_array_equal(orc.normal_stream(300, 1 / 32), _normals(300, 1 / 32))


def test_sign_and_bit_streams_restated(orc):
    s = _uniform_int(1000, 2)
    np.testing.assert_array_equal(orc.sign_stream(1000), np.where(s == 0, -1, 1))
    np.testing.assert_array_equal(orc.bit_stream(1000), _uniform_int(1000, 1))


def test_streams_match_reference_fixture(orc):
    fx = json.load(open(os.path.join(GOLDEN, "rng_streams.json")))
    np.testing.assert_array_equal(orc.normal_stream(64), np.array(fx["normal_stddev1_first64"], np.float32))
    np.testing.assert_array_equal(orc.normal_stream(64, 1 / 32),
                                  np.array(fx["normal_stddev_1_over_32_first64"], np.float32))
    np.testing.assert_array_equal(orc.sign_stream(64), np.array(fx["signs_first64"]))


def test_streams_match_reference_build(orc):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    np.testing.assert_array_equal(orc.normal_stream(5000), orc.ref_random_vectors(5000))
    np.testing.assert_array_equal(orc.sign_stream(5000), orc.ref_random_signs(5000))


# --------------------------------------------------------------- exact k-NN

def _brute(data, queries, k, func):
    d64 = data.astype(np.float64)
    q64 = queries.astype(np.float64)
    if func in (0, 1):
        c = (q64 @ d64.T) / (np.linalg.norm(q64, axis=1)[:, None] * np.linalg.norm(d64, axis=1)[None, :])
        D = -c if func == 0 else 1 - c
    elif func == 2:
        mn = np.minimum(q64[:, None, :], d64[None, :, :]).sum(-1)
        mx = np.maximum(q64[:, None, :], d64[None, :, :]).sum(-1)
        D = 1 - mn / mx
    else:
        D = np.sqrt(((q64[:, None, :] - d64[None, :, :]) ** 2).sum(-1))
    ids = np.stack([np.lexsort((np.arange(len(data)), row))[:k] for row in D])
    return ids, np.take_along_axis(D, ids, 1)


@pytest.mark.parametrize("func", [0, 1, 2, 3])
def test_exact_knn_bruteforce(orc, func):
    rng = np.random.default_rng(func)
    data = rng.standard_normal((700, 9)).astype(np.float32)
    queries = rng.standard_normal((11, 9)).astype(np.float32)
    if func == 2:
        data, queries = np.abs(data), np.abs(queries)
    ids, dist = orc.exact_knn(data, queries, 32, func)
    bids, bd = _brute(data, queries, 32, func)
    np.testing.assert_array_equal(ids, bids)
    np.testing.assert_allclose(dist, bd, rtol=1e-6, atol=1e-7)


def test_exact_knn_index_tie_break_and_padding(orc):
    data = np.array([[1, 0], [1, 0], [0, 1], [1, 0]], dtype=np.float32)
    queries = np.array([[1, 0]], dtype=np.float32)
    ids, dist = orc.exact_knn(data, queries, 32, 1)
    assert list(ids[0, :4]) == [0, 1, 3, 2]
    assert (ids[0, 4:] == -1).all() and (dist[0, 4:] == np.float32(2147483648.0)).all()
    ids0, dist0 = orc.exact_knn(data, queries, 32, 0)
    assert (dist0[0, 4:] == 2.0).all()


def test_exact_knn_matches_reference_goldstandard(orc):
    fx = np.load(os.path.join(GOLDEN, "goldstandard_small.npz"))
    K = int(fx["K"])
    ids, dist = orc.exact_knn(fx["data"], fx["queries"], K, 0)
    np.testing.assert_array_equal(ids, fx["truth_ids"])
    # the truth file prints -cos with 6 significant digits
    np.testing.assert_allclose(dist, fx["truth_dist"], rtol=2e-6, atol=2e-6)


def test_goldstandard_port_matches_reference(orc):
    fx = np.load(os.path.join(GOLDEN, "goldstandard_small.npz"))
    K = int(fx["K"])
    _, ids = orc.goldstandard_port(fx["data"], fx["queries"], K, threads=2)
    np.testing.assert_array_equal(ids, fx["truth_ids"])


# ------------------------------------------------------------- transforms

def test_gj_transform_properties(orc):
    rng = np.random.default_rng(3)
    data = (rng.standard_normal((200, 7)) * 3).astype(np.float32)
    queries = (rng.standard_normal((5, 7)) * 3).astype(np.float32)
    td, tq = orc.gj_transform(data, queries)
    mins = np.minimum(0, np.floor(np.minimum(data.min(0), queries.min(0)))).astype(np.int64)
    shifted = (data + np.abs(mins).astype(np.float32)).astype(np.float64)
    assert (td >= 0).all() and (tq >= 0).all()
    np.testing.assert_allclose(np.linalg.norm(td.astype(np.float64), axis=1), 1.0, atol=3e-7)
    np.testing.assert_allclose(td, shifted / np.linalg.norm(shifted, axis=1)[:, None], rtol=4e-7, atol=1e-8)


# ------------------------------------------------------------------ SimHash

def test_simhash_bits_and_packing(orc):
    rng = np.random.default_rng(11)
    d, words = 20, 3
    x = rng.standard_normal((64, d)).astype(np.float32)
    planes = rng.standard_normal((words * 32, d)).astype(np.float32)
    bits = orc.simhash(x, planes, words)
    dots = x.astype(np.float64) @ planes.astype(np.float64).T  # [n][w*32+b]
    sure = np.abs(dots) > 1e-4 * np.linalg.norm(x, axis=1)[:, None] * np.linalg.norm(planes, axis=1)[None, :]
    unpacked = np.stack([(bits[:, w][:, None] >> np.arange(32)) & 1 for w in range(words)], 1).reshape(64, -1)
    assert sure.mean() > 0.99
    np.testing.assert_array_equal(unpacked[sure], (dots[sure] >= 0).astype(np.uint32))


def test_simhash_generic_bits_per_word(orc):
    # simHash.cuh:20-46 with hashBits = 10: word t holds hyperplanes t*10 .. t*10+9
    rng = np.random.default_rng(12)
    x = rng.standard_normal((50, 8)).astype(np.float32)
    planes = rng.standard_normal((4 * 10, 8)).astype(np.float32)
    keys = orc.simhash(x, planes, 4, 10)
    full = orc.simhash(x, np.concatenate([planes, np.zeros((88, 8), np.float32)]), 4, 32)
    assert (keys < 1024).all()
    flat = np.stack([(keys[:, t][:, None] >> np.arange(10)) & 1 for t in range(4)], 1).reshape(50, -1)
    ref = (x.astype(np.float64) @ planes.T.astype(np.float64) >= 0).astype(np.uint32)
    sure = np.abs(x.astype(np.float64) @ planes.T.astype(np.float64)) > 1e-4
    np.testing.assert_array_equal(flat[sure], ref[sure])
    assert full.shape == (50, 4)


def test_simhash_bits_exact_fma_chain(orc):
    # Every bit, borderline ones included: the sign of the sequential FP32
    # fused multiply-add chain (computed here by an exact emulation of fmaf).
    from _fp32 import dot_fma_chain

    rng = np.random.default_rng(13)
    d, words = 29, 2
    x = rng.standard_normal((80, d)).astype(np.float32)
    planes = rng.standard_normal((words * 32, d)).astype(np.float32)
    dots = dot_fma_chain(x, planes)
    want = (dots >= 0).astype(np.uint32)
    bits = orc.simhash(x, planes, words)
    got = np.stack([(bits[:, w][:, None] >> np.arange(32)) & 1 for w in range(words)], 1).reshape(80, -1)
    np.testing.assert_array_equal(got, want)


# ---------------------------------------------------------------- SimHash-JL

def test_jl_projection_is_the_fma_chain(orc):
    # simHashJL.cuh:34-51: sketch value s of row i = sequential FP32 fmaf chain
    # of x_i . planes[s] in index order, bit for bit; and within the FP32 dot
    # product error bound gamma_d * sum|x||p| of the FP64 product (a swapped
    # operand or a wrong stride would be far outside it).
    from _fp32 import dot_fma_chain

    rng = np.random.default_rng(31)
    n, d, sdim = 60, 37, 24
    x = rng.standard_normal((n, d)).astype(np.float32)
    planes = (rng.standard_normal((sdim, d)) / np.sqrt(sdim)).astype(np.float32)
    got = orc.jl(x, planes, sdim)
    np.testing.assert_array_equal(got, dot_fma_chain(x, planes))
    exact = x.astype(np.float64) @ planes.astype(np.float64).T
    mag = np.abs(x.astype(np.float64)) @ np.abs(planes.astype(np.float64)).T
    gamma = d * 2.0 ** -24 / (1 - d * 2.0 ** -24)
    assert np.all(np.abs(got - exact) <= gamma * mag)


def _jl_sq_l2(ds, q, fused):
    """Squared L2 of every row of ds to q, summed in FP32 in index order:
    fl(t*t) then a separate add (fused=False, the reference's pow + add), or
    one fmaf per term (fused=True)."""
    from _fp32 import fmaf32

    s = np.zeros(ds.shape[0], np.float32)
    for j in range(ds.shape[1]):
        t = (q[j] - ds[:, j]).astype(np.float32)
        s = fmaf32(t, t, s) if fused else (s + t * t).astype(np.float32)
    return s


def test_jl_topc_unfused_squares_and_id_ties(orc):
    # distanceFunctions.cuh:87-99: each term pow(q - x, 2) is fl(t*t), added
    # by a separate FP32 add; candidates ordered by (sum, ID).  The rows are
    # permutations of one vector, so their exact distances to a query at the
    # origin are equal and only the FP32 roundings order them: a fused sum
    # gives a different order (checked below), so the pin tells the two apart.
    rng = np.random.default_rng(32)
    sdim, N, C = 16, 2000, 150
    base = rng.uniform(0.5, 2.0, sdim).astype(np.float32)
    ds = np.stack([rng.permutation(base) for _ in range(N)]).astype(np.float32)
    ds[1500:] = rng.standard_normal((500, sdim)).astype(np.float32) * 2
    qs = np.zeros((4, sdim), np.float32)
    qs[2:] = rng.standard_normal((2, sdim)).astype(np.float32) * 0.01
    ids = orc.jl_topc(ds, qs, C)
    differs = 0
    for q in range(len(qs)):
        s = _jl_sq_l2(ds, qs[q], fused=False)
        order = np.lexsort((np.arange(N), s))[:C]
        np.testing.assert_array_equal(ids[q], order)
        sf = _jl_sq_l2(ds, qs[q], fused=True)
        differs += int(not np.array_equal(np.lexsort((np.arange(N), sf))[:C], order))
    assert differs > 0
    # fewer data rows than C: the tail is padded with -1
    few = orc.jl_topc(ds[:10], qs[:1], 16)
    assert (few[0, 10:] == -1).all() and sorted(few[0, :10]) == list(range(10))


# ------------------------------------------- dense-rotation cross-polytope

def test_bf16_round_matches_torch(orc):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(50)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32),
                        np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 0.0, -0.0], np.float32)])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(orc.bf16_round(x), want)
    G = orc.dense_rotations(2, 2, 8, 5)
    np.testing.assert_array_equal(G.ravel(), orc.bf16_round(orc.normal_stream(2 * 2 * 8 * 5)))


def _dense_hash_rule(y):
    """crossPolyTope.cuh:64-76 over the last axis: first j of largest |y_j|,
    encoded j (negative) or j + P."""
    P = y.shape[-1]
    a = np.abs(y)
    j = np.argmax(a, axis=-1)  # first maximum
    neg = np.take_along_axis(y, j[..., None], -1)[..., 0] < 0
    return np.where(neg, j, j + P).astype(np.uint32)


@pytest.mark.parametrize("d,tables,K", [(16, 3, 2), (12, 2, 3), (40, 2, 1)])
def test_crosspoly_dense_is_the_fma_chain_argmax(orc, d, tables, K):
    # Every hash equals the reference's argmax rule applied to the sequential
    # FP32 fmaf chain of each rotation row (exact emulation), packed K to a
    # key in log2(2P)-bit fields.
    from _fp32 import dot_fma_chain

    rng = np.random.default_rng(51 + d)
    x = rng.standard_normal((70, d)).astype(np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    x[5] = x[4]  # duplicates hash alike
    P = orc.pow2_at_least(d)
    b = int(np.log2(2 * P))
    G = orc.dense_rotations(tables, K, P, d)
    keys = orc.crosspoly_dense(x, tables, K, G)
    for t in range(tables):
        want = np.zeros(len(x), np.uint64)
        for k in range(K):
            y = dot_fma_chain(x, G[t * K + k])
            want |= _dense_hash_rule(y).astype(np.uint64) << np.uint64(k * b)
        np.testing.assert_array_equal(keys[:, t], want.astype(np.uint32))
    assert (keys < (1 << (b * K))).all() if b * K < 32 else True
    np.testing.assert_array_equal(keys[4], keys[5])


def test_crosspoly_dense_agrees_with_fp64_away_from_ties(orc):
    # On hashes whose top two |y| differ by more than twice the FP32 error
    # bound, the argmax of the FP64 product decides (a transposed rotation or a
    # wrong stride would disagree on most of them).
    rng = np.random.default_rng(52)
    d, tables, K = 100, 4, 2
    x = rng.standard_normal((300, d)).astype(np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    P = orc.pow2_at_least(d)
    b = int(np.log2(2 * P))
    G = orc.dense_rotations(tables, K, P, d)
    keys = orc.crosspoly_dense(x, tables, K, G)
    checked = 0
    for h in range(tables * K):
        y = x.astype(np.float64) @ G[h].astype(np.float64).T
        a = np.sort(np.abs(y), axis=1)
        bound = d * 2.0 ** -23 * np.linalg.norm(G[h].astype(np.float64), axis=1).max()
        sure = a[:, -1] - a[:, -2] > 2 * bound
        got = (keys[:, h // K] >> np.uint32((h % K) * b)) & np.uint32((1 << b) - 1)
        np.testing.assert_array_equal(got[sure], _dense_hash_rule(y)[sure])
        checked += int(sure.sum())
    assert checked > 0.95 * 300 * tables * K


def _popcount(a):
    a = a.astype(np.uint32)
    return np.unpackbits(a.view(np.uint8).reshape(a.shape + (4,)), axis=-1).sum(-1).sum(-1)


@pytest.mark.parametrize("kind", [0, 1])
def test_sketch_topc_bruteforce(orc, kind):
    rng = np.random.default_rng(5 + kind)
    hi = 2 ** 32 if kind == 0 else 3  # few distinct values -> many ties
    ds = rng.integers(0, hi, size=(3000, 6), dtype=np.uint64).astype(np.uint32)
    qs = rng.integers(0, hi, size=(7, 6), dtype=np.uint64).astype(np.uint32)
    ids, dist = orc.sketch_topc(kind, ds, qs, 100)
    for q in range(7):
        if kind == 0:
            h = _popcount(ds ^ qs[q])
        else:
            h = (ds != qs[q]).sum(1)
        order = np.lexsort((np.arange(3000), h))[:100]
        np.testing.assert_array_equal(ids[q], order)
        np.testing.assert_array_equal(dist[q], h[order])


def test_rerank_bruteforce(orc):
    rng = np.random.default_rng(9)
    data = rng.standard_normal((400, 10)).astype(np.float32)
    queries = rng.standard_normal((6, 10)).astype(np.float32)
    cand = np.stack([rng.permutation(400)[:90] for _ in range(6)]).astype(np.int32)
    cand[0, 10:20] = -1
    ids, dist = orc.rerank(data, queries, cand, 32, 1)
    for q in range(6):
        c = cand[q][cand[q] >= 0]
        bids, bd = _brute(data[c], queries[q:q + 1], 32, 1)
        np.testing.assert_array_equal(ids[q], c[bids[0]])


# --------------------------------------------------------------- validation

def test_recall_and_ratio_match_reference_fixture(orc):
    fx = json.load(open(os.path.join(GOLDEN, "validation_small.json")))
    t = np.array(fx["truth_ids"], np.int32)
    r = np.array(fx["result_ids"], np.int32)
    rd = np.array(fx["result_dist"], np.float32)
    td = np.array(fx["truth_dist"], np.float32)
    assert orc.recall(t, r, fx["report_k"]) == pytest.approx(fx["recall"], abs=1e-6)
    assert orc.distance_ratio(td, r, rd, fx["report_k"]) == pytest.approx(fx["distance_ratio"], rel=1e-5)


def test_recall_closed_form(orc):
    truth = np.arange(64, dtype=np.int32).reshape(2, 32)
    res = truth.copy()
    res[1, :8] = 1000 + np.arange(8)
    assert orc.recall(truth, res, 32) == pytest.approx((1.0 + 24 / 32) / 2)


# ---------------------------------------------------------- cross-polytope

def test_crosspoly_matches_hadamard(orc):
    from scipy.linalg import hadamard

    rng = np.random.default_rng(21)
    d, tables, P = 12, 5, 16
    x = orc.normalize_rows(rng.standard_normal((300, d)).astype(np.float32))
    signs = np.where(rng.random(tables * 3 * P) < 1 / 3, -1, 1).astype(np.int32)
    keys = orc.crosspoly(x, tables, signs)
    H = hadamard(P).astype(np.float64)
    ok = 0
    for t in range(tables):
        v = np.zeros((300, P))
        v[:, :d] = x
        for r in range(3):
            v = (v * signs[(t * 3 + r) * P:(t * 3 + r + 1) * P]) @ H.T
        a = np.abs(v)
        top = np.argsort(-a, axis=1, kind="stable")
        gap = a[np.arange(300), top[:, 0]] - a[np.arange(300), top[:, 1]]
        i = top[:, 0]
        expect = np.where(v[np.arange(300), i] < 0, i, i + P)
        sure = gap > 1e-4
        np.testing.assert_array_equal(keys[sure, t], expect[sure])
        ok += sure.sum()
    assert ok > 0.98 * 300 * tables


def test_hadamard_involution():
    from scipy.linalg import hadamard

    H = hadamard(128)
    np.testing.assert_array_equal(H @ H, 128 * np.eye(128, dtype=H.dtype))


# -------------------------------------------------------- weighted MinHash

@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_known_answers(orc, ctr, key, expect):
    # Random123 known-answer vectors for Philox4x32-10
    np.testing.assert_array_equal(orc.philox4x32_10(ctr, key), np.array(expect, np.uint32))


def test_wmh_collision_probability_tracks_generalized_jaccard(orc):
    rng = np.random.default_rng(31)
    d, H = 16, 2048
    base = rng.standard_exponential(d).astype(np.float32)
    pairs = []
    for s in (0.05, 0.3, 0.8):
        other = np.abs(base + s * rng.standard_normal(d)).astype(np.float32)
        pairs.append(other)
    x = orc.normalize_rows(np.stack([base] + pairs))
    bounds, start, M = orc.wmh_bounds(x)
    assert M == int(np.ceil(x.max(0)).clip(min=0).sum())
    sk = orc.wmh_sketch(x, bounds, start, M, H, 0, np.zeros(1, np.uint8))
    for i in range(1, 4):
        gj = np.minimum(x[0], x[i]).sum() / np.maximum(x[0], x[i]).sum()
        est = (sk[0] == sk[i]).mean()
        assert abs(est - gj) < 4 * math.sqrt(gj * (1 - gj) / H) + 0.01


def test_wmh_counts_are_geometric(orc):
    rng = np.random.default_rng(32)
    x = orc.normalize_rows(rng.standard_exponential((40, 12)).astype(np.float32))
    bounds, start, M = orc.wmh_bounds(x)
    sk = orc.wmh_sketch(x, bounds, start, M, 400, 0, np.zeros(1, np.uint8))
    # P(green) = sum(x)/M, so the mean number of reds is M/sum(x) - 1
    expect = M / x.sum(1) - 1
    np.testing.assert_allclose(sk.mean(1), expect, rtol=0.25)


# ---------------------------------------------------------------------- LSH

def test_lsh_candidates_bruteforce(orc):
    rng = np.random.default_rng(41)
    dk = rng.integers(0, 4, size=(500, 3)).astype(np.uint32)
    qk = rng.integers(0, 4, size=(2, 3)).astype(np.uint32)
    for q in range(2):
        c = orc.lsh_candidates(dk, qk, q)
        expect = np.nonzero((dk == qk[q]).any(1))[0]
        np.testing.assert_array_equal(c, expect)


def test_wmh_one_bit_words_pack_the_counts(orc):
    """Mode 1 word w, bit b is bitmap[count of hash w*hbits + b] (the counts
    themselves are pinned above), for the 32-bit words of the linear scan and
    the bucketKeyBits-wide words of the LSH keys."""
    rng = np.random.default_rng(33)
    x = orc.normalize_rows(rng.standard_exponential((30, 10)).astype(np.float32))
    bounds, start, M = orc.wmh_bounds(x)
    bm = orc.bit_stream(4096)
    for hbits, words in ((32, 2), (7, 5)):
        counts = orc.wmh_sketch(x, bounds, start, M, words * hbits, 0, bm)
        packed = orc.wmh_sketch(x, bounds, start, M, words, 1, bm, hbits=hbits)
        bits = bm[counts % 4096].astype(np.uint64).reshape(30, words, hbits)
        expect = (bits << np.arange(hbits, dtype=np.uint64)).sum(2).astype(np.uint32)
        np.testing.assert_array_equal(packed, expect)


def _angular64(x, q):
    x = x.astype(np.float64)
    q = q.astype(np.float64)
    return 1.0 - (x @ q) / (np.linalg.norm(x, axis=1) * np.linalg.norm(q))


def test_lsh_search_bruteforce_exact(orc):
    """Exact mode: the k best of the points sharing a key with the query in
    at least one table, by (distance, ID); padding when fewer exist."""
    rng = np.random.default_rng(42)
    N, Q, d, L, bits = 800, 6, 12, 3, 6
    data = rng.standard_normal((N, d)).astype(np.float32)
    queries = rng.standard_normal((Q, d)).astype(np.float32)
    ids, dist = orc.lsh_search(data, queries, 40, 2, 3, 1, 1, bits, L, False)
    assert ids.shape == (Q, 64)
    dk = orc.lsh_keys(data, 3, L, bits)
    qk = orc.lsh_keys(queries, 3, L, bits)
    planes = orc.normal_stream(L * bits * d).reshape(L * bits, d)
    proj = (data.astype(np.float64) @ planes.T.astype(np.float64)) >= 0
    weights = (1 << np.arange(bits)).astype(np.uint32)
    approx = (proj.reshape(N, L, bits) * weights).sum(2)
    assert (approx != dk).mean() < 0.01  # the FP32 bits agree with FP64 signs away from zero
    for q in range(Q):
        cand = np.nonzero((dk == qk[q]).any(1))[0]
        dd = _angular64(data[cand], queries[q])
        order = np.lexsort((cand, dd))[:64]
        want = np.full(64, -1)
        want[:order.size] = cand[order]
        np.testing.assert_array_equal(ids[q], want)
        np.testing.assert_allclose(dist[q, :order.size], dd[order], rtol=1e-6, atol=1e-7)
        assert np.all(dist[q, order.size:] == np.float32(2147483648.0))


def test_lsh_search_bruteforce_sketched(orc):
    """Sketched mode: the top C of the candidate union by (Hamming distance,
    ID), reranked exactly."""
    rng = np.random.default_rng(43)
    N, Q, d, L, bits, W, C = 600, 5, 16, 4, 4, 2, 25
    data = rng.standard_normal((N, d)).astype(np.float32)
    queries = rng.standard_normal((Q, d)).astype(np.float32)
    ids, _ = orc.lsh_search(data, queries, 10, 3, 3, W, 1, bits, L, True, C)
    dk, qk = orc.lsh_keys(data, 3, L, bits), orc.lsh_keys(queries, 3, L, bits)
    planes = orc.normal_stream(W * 32 * d)
    ds, qs = orc.simhash(data, planes, W), orc.simhash(queries, planes, W)
    for q in range(Q):
        cand = np.nonzero((dk == qk[q]).any(1))[0]
        ham = np.array([sum(bin(int(a) ^ int(b)).count("1") for a, b in zip(ds[i], qs[q])) for i in cand])
        top = cand[np.lexsort((cand, ham))[:C]]
        dd = _angular64(data[top], queries[q])
        order = np.lexsort((top, dd))[:32]
        want = np.full(32, -1)
        want[:order.size] = top[order]
        np.testing.assert_array_equal(ids[q], want)


def test_bucket_distribution_pins(orc):
    # Hand case (255 has its own bin) and the library histogram.
    c = orc.bucket_distribution(np.array([255, 255, 0, 7], np.uint8))
    assert c[255] == 2 and c[0] == 1 and c[7] == 1 and c.sum() == 4
    h = np.random.default_rng(5).integers(0, 256, 10007).astype(np.uint8)
    np.testing.assert_array_equal(orc.bucket_distribution(h), np.bincount(h, minlength=256))


def test_bucket_distribution_full_key_pins(orc):
    # The first byte decides the order (bit strings compare as the bytes do).
    k, c = orc.bucket_distribution_full_key(np.array([[1, 0], [0, 255], [1, 0]], np.uint8), 2)
    np.testing.assert_array_equal(k, [[0, 255], [1, 0]])
    np.testing.assert_array_equal(c, [1, 2])
    rng = np.random.default_rng(6)
    for sd in (1, 3, 9):
        pool = rng.integers(0, 256, (40, sd)).astype(np.uint8)
        h = pool[rng.integers(0, 40, 2000)]
        k, c = orc.bucket_distribution_full_key(h, sd)
        uk, uc = np.unique(h, axis=0, return_counts=True)
        np.testing.assert_array_equal(k, uk)
        np.testing.assert_array_equal(c, uc)


def test_bucket_occupancy_pins(orc):
    s, h = orc.bucket_occupancy(np.array([3, 3, 3, 0, 5]), 8, 3)
    assert s == {"buckets": 8, "nonempty": 3, "points": 5, "max_size": 3}
    np.testing.assert_array_equal(h, [5, 2, 1])  # sizes 0 x5, 1 x2, >=2 x1
    keys = np.random.default_rng(7).integers(0, 1024, 5000)
    sizes = np.bincount(keys, minlength=1024)
    s, h = orc.bucket_occupancy(keys, 1024, 12)
    assert s["nonempty"] == np.count_nonzero(sizes) and s["max_size"] == sizes.max() and s["points"] == 5000
    np.testing.assert_array_equal(h, np.bincount(np.minimum(sizes, 11), minlength=12))
