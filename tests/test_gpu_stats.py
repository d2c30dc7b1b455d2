"""Bucket and sketch statistics on the GPU against the oracle:
statistics.cuh:10-16 (byte histogram), statisticsCpu.cpp:11-28 (full-key
distribution) and the occupancy of (K,L) index tables."""
import numpy as np
import pytest


@pytest.mark.gpu
@pytest.mark.parametrize("n", [0, 1, 15, 16, 1_000_003])
def test_bucket_distribution(lsh, orc, n):
    rng = np.random.default_rng(n)
    h = np.minimum(rng.geometric(0.05, n), 255).astype(np.uint8)  # skewed, with 255 present at scale
    got = lsh.bucket_distribution(h)
    np.testing.assert_array_equal(got, orc.bucket_distribution(h))
    assert got.sum() == n


@pytest.mark.gpu
def test_bucket_distribution_of_wmh_byte_sketches(lsh, orc):
    # weightedMinHash.cuh:215-229 runs it over the byte sketches of the data.
    x = np.abs(np.random.default_rng(3).standard_normal((3000, 20))).astype(np.float32)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    sk = lsh.wmh_sketch(x, 8).astype(np.uint8)
    np.testing.assert_array_equal(lsh.bucket_distribution(sk), orc.bucket_distribution(sk))


@pytest.mark.gpu
@pytest.mark.parametrize("sd,n,pool", [(1, 5000, 300), (3, 20000, 500), (8, 30000, 2000), (9, 20000, 700),
                                       (17, 4000, 4000), (2, 1, 1)])
def test_bucket_distribution_full_key(lsh, orc, sd, n, pool):
    rng = np.random.default_rng(sd * 1000 + n)
    p = rng.integers(0, 256, (pool, sd)).astype(np.uint8)
    p[: pool // 3, : sd // 2] = 0  # shared prefixes so later words decide
    h = p[rng.integers(0, pool, n)]
    k, c = lsh.bucket_distribution_full_key(h)
    ok, oc = orc.bucket_distribution_full_key(h, sd)
    np.testing.assert_array_equal(k, ok)
    np.testing.assert_array_equal(c, oc)
    assert c.sum() == n


@pytest.mark.gpu
def test_bucket_distribution_full_key_empty(lsh):
    k, c = lsh.bucket_distribution_full_key(np.zeros((0, 4), np.uint8))
    assert k.shape == (0, 4) and c.size == 0


@pytest.mark.gpu
@pytest.mark.parametrize("keys,bits,tables,hist_len", [(3, 10, 4, 0), (3, 14, 3, 40), (7, 6, 5, 20000),
                                                      (8, 12, 3, 50), (3, 20, 2, 30)])
def test_index_bucket_stats(lsh, orc, keys, bits, tables, hist_len):
    # (8, 12): dense-rotation keys in sorted tables; (3, 20): SimHash keys
    # wider than 16 bits, also sorted (the 16-bit key type keeps 2^16 buckets).
    rng = np.random.default_rng(bits)
    data = rng.standard_normal((30000, 24)).astype(np.float32)
    with lsh.Index(data, implementation=2, framework=1, keysImplementation=keys, bucketKeyBits=bits,
                   tables=tables) as ix:
        x = orc.normalize_rows(data) if keys in (7, 8) else data
        dk = orc.lsh_keys(x, keys, tables, bits)
        buckets = 2 ** (min(bits, 16) if keys == 3 else bits)
        for t in range(tables):
            s, h = ix.bucket_stats(t, hist_len)
            os_, oh = orc.bucket_occupancy(dk[:, t], buckets, hist_len)
            assert s == os_
            np.testing.assert_array_equal(h, oh)
        with pytest.raises(lsh.LshError):
            ix.bucket_stats(tables, 4)


@pytest.mark.gpu
def test_index_bucket_stats_needs_buckets(lsh):
    data = np.random.default_rng(1).standard_normal((100, 8)).astype(np.float32)
    with lsh.Index(data, implementation=2) as ix:
        with pytest.raises(lsh.LshError):
            ix.bucket_stats(0)
