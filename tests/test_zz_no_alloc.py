"""lsh_hamming_topc_noalloc (device pointers, working memory from the
library's per-GPU scratch array) against lsh_hamming_topc, called through the
C ABI: two calls in a row, the second smaller, and a request larger than the
scratch array, which is refused before any work is queued."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sketches(rng, n, words):
    centres = rng.integers(0, 2 ** 32, size=(40, words), dtype=np.uint64).astype(np.uint32)
    flips = np.packbits((rng.random((n, words * 32)) < 0.15).astype(np.uint8), axis=1).view(np.uint32)
    return (centres[rng.integers(0, 40, n)] ^ flips).astype(np.uint32)


def _noalloc(lsh, torch, ds, qs, C_):
    dev = torch.device("cuda:0")
    d_ds = torch.from_numpy(ds.view(np.int32)).to(dev)
    d_qs = torch.from_numpy(qs.view(np.int32)).to(dev)
    ids = torch.full((len(qs), C_), 7, dtype=torch.int32, device=dev)
    dist = torch.full((len(qs), C_), 7, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    st = lsh.load().lsh_hamming_topc_noalloc(C.c_void_p(d_ds.data_ptr()), len(ds), C.c_void_p(d_qs.data_ptr()),
                                              len(qs), ds.shape[1], C_, C.c_void_p(ids.data_ptr()),
                                              C.c_void_p(dist.data_ptr()), C.c_void_p(stream.cuda_stream))
    assert st == 0, lsh.load().lsh_last_error()
    torch.cuda.synchronize()
    return ids.cpu().numpy(), dist.cpu().numpy()


def test_noalloc_topc_matches_allocating_entry_point(lsh):
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2024)
    # 12 words: data rows not a multiple of 8 words (read word by word)
    ds, qs = _sketches(rng, 30000, 12), _sketches(rng, 300, 12)
    qs[:5] = ds[:5]
    want_ids, want_d = lsh.hamming_topc(ds, qs, 200)
    ids, dist = _noalloc(lsh, torch, ds, qs, 200)
    np.testing.assert_array_equal(ids, want_ids)
    np.testing.assert_array_equal(dist, want_d)
    # second call, smaller: 8 words, fewer points than C for some rows' ties,
    # the previous call's state still in the array
    ds2, qs2 = _sketches(rng, 7000, 8), _sketches(rng, 37, 8)
    want_ids, want_d = lsh.hamming_topc(ds2, qs2, 50)
    ids, dist = _noalloc(lsh, torch, ds2, qs2, 50)
    np.testing.assert_array_equal(ids, want_ids)
    np.testing.assert_array_equal(dist, want_d)
    # fewer points than C: padded with -1
    ids, dist = _noalloc(lsh, torch, ds2[:20], qs2[:3], 32)
    want_ids, want_d = lsh.hamming_topc(ds2[:20], qs2[:3], 32)
    np.testing.assert_array_equal(ids, want_ids)
    assert (ids[:, 20:] == -1).all() and (dist[:, 20:] == -1).all()


def test_noalloc_topc_rejects_oversized_request(lsh):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    buf = torch.zeros(16, dtype=torch.int32, device=dev)
    p = C.c_void_p(buf.data_ptr())
    # 4096 candidates for 2,000 queries needs about 2000 * 8 * 8704 bytes
    st = lsh.load().lsh_hamming_topc_noalloc(p, 1, p, 2000, 8, 4096, p, p, None)
    assert st == 4
    assert b"allocation-free scratch" in lsh.load().lsh_last_error()
