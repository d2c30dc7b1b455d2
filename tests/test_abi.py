"""The C ABI boundary without a GPU: the library loads, exports every entry
point include/lshgpu.h declares, and the host-only entry points and argument
checks behave like the reference surface."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "lshgpu.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lsh_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_reference_surface():
    names = declared_functions()
    for n in ("lsh_linear_scans", "lsh_framework", "lsh_calculate_k", "lsh_result_free", "lsh_last_error",
              "lsh_index_create", "lsh_index_search", "lsh_index_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol(lsh):
    lib = lsh.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) == set(lsh.EXPORTED_SYMBOLS)


def test_struct_layouts_match_reference():
    import lshongpu_amd as m

    assert C.sizeof(m._Point) == 8  # point.h: int + float
    assert m._Result.results.offset == 0
    assert C.sizeof(m._Result) == 8 + 3 * 8 + 4 + 4  # LP64 clock_t is long
    assert m.POINT_DTYPE.itemsize == 8


def test_calculate_k_matches_oracle(lsh, orc):
    for k in (1, 5, 32, 33, 100, 1000, 1024, 4000):
        assert lsh.calculateK(k) == orc.calculate_k(k)


def test_random_streams_match_oracle_and_fixture(lsh, orc):
    np.testing.assert_array_equal(lsh.random_normals(4096), orc.normal_stream(4096))
    np.testing.assert_array_equal(lsh.random_normals(100, 1 / 16), orc.normal_stream(100, 1 / 16))
    np.testing.assert_array_equal(lsh.random_signs(999), orc.sign_stream(999))


@pytest.mark.parametrize("impl", [0, 7, -1])
def test_invalid_implementation_is_an_error_not_an_exit(lsh, impl):
    x = np.zeros((4, 3), np.float32)
    with pytest.raises(lsh.LshError) as e:
        lsh.linearScans(impl, 32, 3, 4, 4, x, x, 1, 1)
    assert e.value.code == 1
    assert "implementation" in e.value.message


def test_invalid_distance_function(lsh):
    x = np.zeros((4, 3), np.float32)
    with pytest.raises(lsh.LshError) as e:
        lsh.linearScans(2, 32, 3, 4, 4, x, x, 1, 9)
    assert e.value.code == 1 and "distance" in e.value.message


def test_invalid_lsh_keys(lsh):
    x = np.zeros((4, 3), np.float32)
    with pytest.raises(lsh.LshError) as e:
        lsh.LSH(3, 6, 32, 3, 4, 4, x, x, 1, 1, 8, 4, True)
    assert e.value.code == 1 and "keys" in e.value.message


def test_null_out_pointer(lsh):
    lib = lsh.load()
    st = lib.lsh_linear_scans(2, 32, 3, 0, 0, None, None, 1, 1, None)
    assert st == 1
    assert b"out" in lib.lsh_last_error()


def test_short_buffers_rejected(lsh):
    x = np.zeros((4, 3), np.float32)
    with pytest.raises(lsh.LshError):
        lsh.linearScans(2, 32, 3, 10, 4, x, x, 1, 1)


def _index_shell(lsh, d):
    """An Index object whose argument checks run without a device (its
    handle is null, so nothing reaches the library)."""
    ix = lsh.Index.__new__(lsh.Index)
    ix.d, ix.n, ix._h = d, 0, C.c_void_p()
    return ix


def test_index_search_rejects_bad_host_buffers(lsh):
    ix = _index_shell(lsh, 8)
    good = np.zeros((3, 8), np.float32)
    with pytest.raises(lsh.LshError, match="2-D"):
        ix.search(np.zeros(8, np.float32), 32)
    with pytest.raises(lsh.LshError, match="d = 8"):
        ix.search(np.zeros((3, 7), np.float32), 32)
    with pytest.raises(lsh.LshError, match="k must be"):
        ix.search(good, 0)
    short = np.empty((3, 31), dtype=lsh.POINT_DTYPE)
    with pytest.raises(lsh.LshError, match="need"):
        ix.search(good, 32, out=short)
    strided = np.empty((3, 64), dtype=lsh.POINT_DTYPE)[:, ::2]
    with pytest.raises(lsh.LshError, match="contiguous"):
        ix.search(good, 32, out=strided)


def test_index_search_rejects_bad_tensors(lsh):
    torch = pytest.importorskip("torch")
    ix = _index_shell(lsh, 8)
    good = torch.zeros((3, 8), dtype=torch.float32)
    with pytest.raises(lsh.LshError, match="float32"):
        ix.search(torch.zeros((3, 8), dtype=torch.float16), 32)
    with pytest.raises(lsh.LshError, match="contiguous"):
        ix.search(torch.zeros((8, 3), dtype=torch.float32).t(), 32)
    with pytest.raises(lsh.LshError, match="d = 8"):
        ix.search(torch.zeros((3, 9), dtype=torch.float32), 32)
    with pytest.raises(lsh.LshError, match="need"):
        ix.search(good, 32, out=torch.empty((3, 63), dtype=torch.int32))
    with pytest.raises(lsh.LshError, match="float32"):
        lsh.Index(torch.zeros((16, 8), dtype=torch.float64))
