"""The command-line runners on the GPU: knn (kernel.cu:119-196, same 15
arguments) and goldstandard (truth files, goldstandard.cpp:175-200)."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "lshongpu_amd")


def _run(args):
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def _unit(rng, n, d):
    x = rng.standard_normal((n, d)).astype(np.float32)
    return x / np.linalg.norm(x, axis=1, keepdims=True)


@pytest.mark.gpu
def test_goldstandard_cli_matches_oracle(lsh, orc, tmp_path):
    rng = np.random.default_rng(90)
    data, queries = _unit(rng, 3000, 24), _unit(rng, 20, 24)
    lsh.write_text_dataset(tmp_path / "d.txt", data)
    lsh.write_text_dataset(tmp_path / "q.txt", queries)
    for func in (0, 1, 3):
        out = tmp_path / f"truth{func}.txt"
        _run([os.path.join(PKG, "goldstandard"), "10", str(tmp_path / "d.txt"), str(tmp_path / "q.txt"), str(out),
              str(func)])
        ids, dist = lsh.read_truth(out)
        oids, odist = orc.exact_knn(data, queries, 10, func)
        np.testing.assert_array_equal(ids, oids)
        np.testing.assert_allclose(dist, odist, rtol=1e-5, atol=1e-6)


@pytest.mark.gpu
def test_knn_cli_runs_validates_and_reports(lsh, tmp_path):
    rng = np.random.default_rng(91)
    data, queries = _unit(rng, 20000, 32), _unit(rng, 50, 32)
    d, q, t, csv = (str(tmp_path / n) for n in ("d.txt", "q.txt", "truth.txt", "perf.csv"))
    lsh.write_text_dataset(d, data)
    lsh.write_text_dataset(q, queries)
    _run([os.path.join(PKG, "goldstandard"), "32", d, q, t])
    knn = os.path.join(PKG, "knn")
    # Exact scan: every neighbour of the -cos truth is found.
    out = _run([knn, d, q, t, "1", "1", "32", "2", "0", "1", "0", "0", "0", "0", "0", csv])
    assert "Averge recall: 1" in out
    lines = out.splitlines()
    at = lines.index("50")
    assert lines[at + 1] == "32" and lines[at + 2] == "0:"
    ref = lsh.linearScans(2, 32, 32, 50, 20000, data, queries, 0, 1)
    got = [int(l.split()[0]) for l in lines[at + 3:at + 35]]
    assert got == list(ref.ids[0])
    # SimHash sketch scan and the (K,L) framework through the same runner.
    _run([knn, d, q, t, "1", "0", "32", "3", "8", "1", "0", "0", "0", "0", "0", csv])
    _run([knn, d, q, t, "1", "0", "32", "2", "0", "1", "1", "12", "10", "3", "0", csv])
    rows = [l.split(",") for l in open(csv).read().splitlines()]
    assert [len(r) for r in rows] == [13, 13, 13]
    assert rows[0][:8] == ["2", "0", "0", "32", "128", "0", "0", "0"]
    assert rows[2][:8] == ["2", "3", "0", "32", "128", "12", "10", "0"]
    assert float(rows[0][11]) == 1.0 and 0.3 < float(rows[1][11]) <= 1.0
    # The runner's LSH recall is the library's on the same (exactly re-read) data.
    api = lsh.LSH(2, 3, 32, 32, 50, 20000, data, queries, 0, 1, 12, 10, 0)
    recall, _ = lsh.validate(t, api, 32)
    assert float(rows[2][11]) == pytest.approx(recall, abs=1e-6) and recall > 0.05
