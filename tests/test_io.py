"""Datasets, truth files, validation and report formats (host code, no GPU).

Formats follow the reference's runner: gloveparser.cu:9-86 and
DataParsing/Writer/writer.cpp:10-29 (text datasets), siftParser.cpp:8-28
(.fvecs), goldstandard.cpp:123-134 and validation.cpp:62-107 (truth files),
validation.cpp:129-208 (recall, distance ratio), resultWriter.cpp:7-60
(result dump, 13-column CSV).
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN


def test_text_dataset_round_trip(lsh, tmp_path):
    x = np.random.default_rng(1).standard_normal((17, 5)).astype(np.float32) * 1e3
    p = tmp_path / "x.txt"
    lsh.write_text_dataset(p, x)
    y, ids = lsh.read_text_dataset(p)
    np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(ids, np.arange(17))
    lines = p.read_text().splitlines()
    assert lines[:2] == ["17", "5"] and lines[2].startswith("0 ") and lines[2].endswith(" ")


def test_text_dataset_reads_the_oracle_writer_and_hard_rows(lsh, orc, tmp_path):
    x = np.random.default_rng(2).standard_normal((9, 4)).astype(np.float32)
    p = tmp_path / "o.txt"
    orc.write_text_dataset(str(p), x)
    np.testing.assert_array_equal(lsh.read_text_dataset(p)[0], x)
    # Integer parts with several digits, exponents, a row without the
    # trailing space, blank lines and CRLF: the reference's digit loop keeps
    # only the last digit of "12.5" and drops "7" at the end of the row.
    q = tmp_path / "h.txt"
    q.write_text("3\n2\n10 12.5 -3e-2 \n\n11 -0.25 7\r\n12 1E2 +4.0 \n")
    y, ids = lsh.read_text_dataset(q)
    np.testing.assert_array_equal(y, np.array([[12.5, -0.03], [-0.25, 7], [100, 4]], np.float32))
    np.testing.assert_array_equal(ids, [10, 11, 12])


@pytest.mark.parametrize("text,msg", [("2\n2\n0 1 2 \n", "rows"), ("1\n3\n0 1 2 \n", "values"),
                                      ("", "header")])
def test_text_dataset_errors(lsh, tmp_path, text, msg):
    p = tmp_path / "bad.txt"
    p.write_text(text)
    with pytest.raises(lsh.LshError) as e:
        lsh.read_text_dataset(p)
    assert e.value.code == 6 and msg in e.value.message
    with pytest.raises(lsh.LshError) as e:
        lsh.read_text_dataset(tmp_path / "missing.txt")
    assert e.value.code == 6


def test_fvecs(lsh, tmp_path):
    x = np.random.default_rng(3).standard_normal((6, 7)).astype(np.float32)
    rec = np.empty((6, 8), dtype=np.float32)
    rec[:, 0] = np.frombuffer(np.full(6, 7, np.int32).tobytes(), np.float32)
    rec[:, 1:] = x
    p = tmp_path / "x.fvecs"
    rec.tofile(p)
    np.testing.assert_array_equal(lsh.read_fvecs(p), x)
    with open(p, "ab") as f:
        f.write(np.int32(7).tobytes() + x[0, :3].tobytes())  # truncated record
    with pytest.raises(lsh.LshError):
        lsh.read_fvecs(p)


def _points(ids, dist):
    pts = np.empty(ids.shape, dtype=[("ID", np.int32), ("distance", np.float32)])
    pts["ID"], pts["distance"] = ids, dist
    return pts


def test_truth_file_round_trip(lsh, orc, tmp_path):
    gs = np.load(os.path.join(GOLDEN, "goldstandard_small.npz"))
    ids, dist = gs["truth_ids"], gs["truth_dist"]
    pts = _points(np.pad(ids, ((0, 0), (0, 12))), np.pad(dist, ((0, 0), (0, 12))))
    p = tmp_path / "t.txt"
    lsh.write_truth(p, pts, ids.shape[1], query_ids=np.arange(100, 100 + ids.shape[0]))
    tids, tdist = lsh.read_truth(p)
    np.testing.assert_array_equal(tids, ids)
    np.testing.assert_allclose(tdist, dist, rtol=1e-5)  # six significant digits, as goldstandard.cpp writes
    oids, odist = orc.read_truth_file(str(p))
    np.testing.assert_array_equal(oids, tids)
    np.testing.assert_array_equal(odist, tdist)
    assert p.read_text().splitlines()[2] == "100:"


def test_validate_matches_reference_fixture(lsh, tmp_path):
    fx = json.load(open(os.path.join(GOLDEN, "validation_small.json")))
    t = np.array(fx["truth_ids"], np.int32)
    td = np.array(fx["truth_dist"], np.float32)
    p = tmp_path / "truth.txt"
    with open(p, "w") as f:  # the file the reference's runValidation read
        f.write(f"{t.shape[0]}\n{t.shape[1]}\n")
        for q in range(t.shape[0]):
            f.write(f"{q}:\n")
            for j in range(t.shape[1]):
                f.write(f"{t[q, j]} {repr(float(td[q, j]))}\n")
    pts = _points(np.array(fx["result_ids"], np.int32), np.array(fx["result_dist"], np.float32))
    recall, ratio = lsh.validate(p, pts, fx["report_k"])
    assert recall == pytest.approx(fx["recall"], abs=0)
    assert ratio == pytest.approx(fx["distance_ratio"], rel=1e-7)


def test_validate_matches_oracle_on_random_cases(lsh, orc, tmp_path):
    rng = np.random.default_rng(4)
    Q, KT, k, rk = 7, 50, 64, 40
    t = np.stack([rng.permutation(300)[:KT] for _ in range(Q)]).astype(np.int32)
    td = np.sort(rng.random((Q, KT)).astype(np.float32) + 0.1, axis=1)
    r = rng.integers(0, 300, (Q, k)).astype(np.int32)
    rd = rng.random((Q, k)).astype(np.float32)
    p = tmp_path / "truth.txt"
    lsh.write_truth(p, _points(t, td), KT)
    tt, ttd = lsh.read_truth(p)
    recall, ratio = lsh.validate(p, _points(r, rd), rk)
    assert recall == pytest.approx(orc.recall(tt, r, rk), abs=1e-6)
    assert ratio == pytest.approx(orc.distance_ratio(ttd, r, rd, rk), rel=1e-5)


def test_validate_rejects_short_results(lsh, tmp_path):
    p = tmp_path / "truth.txt"
    p.write_text("3\n1\n0:\n1 0.5\n1:\n2 0.5\n2:\n3 0.5\n")
    with pytest.raises(lsh.LshError):
        lsh.validate(p, _points(np.zeros((2, 32), np.int32), np.zeros((2, 32), np.float32)), 1)


def test_write_results_format(lsh, tmp_path):
    pts = _points(np.array([[3, 1, 9], [4, -1, 2]], np.int32), np.array([[0.5, 1.25, 2], [0, 2147483648.0, 1]],
                                                                         np.float32))
    p = tmp_path / "r.txt"
    lsh.write_results(p, pts, 2)
    assert p.read_text() == "2\n2\n0:\n3 0.500000\n1 1.250000\n1:\n4 0.000000\n-1 2147483648.000000\n"


def test_write_performance_csv(lsh, tmp_path):
    p = tmp_path / "perf.csv"
    r = lsh.Result(np.zeros((0, 32), dtype=lsh.POINT_DTYPE), 3, 120, 45, 0.953125, 1.0421)
    lsh.write_performance(p, r, 3, 0, 32, 32, 0, 0)
    lsh.write_performance(p, r, 2, 7, 1, 64, 8, 16)
    lines = p.read_text().splitlines()
    assert lines == ["3,0,32,32,128,0,0,0,3,120,45,0.953125,1.0421", "2,7,1,64,128,8,16,0,3,120,45,0.953125,1.0421"]
    assert all(len(l.split(",")) == 13 for l in lines)
