"""Writes the golden fixtures of tests/golden/ from the reference's own code.

Run in a container that has the reference sources (oracle/_ref/ is built from
them by `make -C oracle ref`):

    python tests/golden/make_golden.py

Fixtures (inputs and expected outputs, nothing regenerated at test time):
  rng_streams.json      first values of the reference's host random streams
                        (randomVectorGenerator.cpp:32-87), from oracle/_ref/libref.so
  goldstandard_small.npz  a small dataset, its queries and the truth file the
                        reference's GoldStandard program writes for them
                        (goldstandard.cpp:136-173), from oracle/_ref/goldstandard
  validation_small.json recall and distance ratio the reference's
                        runValidation (validation.cpp:201-208) reports for a
                        given truth file and result array
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402


def rng_streams() -> dict:
    # The (N, sketchDim) variant keeps a function-static buffer sized by its
    # first call, so it runs in a child process of its own.
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, "sd32.json")
        code = (
            "import sys, json; sys.path.insert(0, %r); import oracle, numpy as np;"
            "v = np.empty(64, dtype=np.float32); oracle._ref().ref_random_vectors_sd("
            "v.ctypes.data_as(oracle._FP), 64, 32); json.dump([float(x) for x in v], open(%r, 'w'))"
            % (ROOT, out)
        )
        subprocess.run([sys.executable, "-c", code], capture_output=True, check=True)
        sd32 = json.load(open(out))
    return {
        "source": "LinearScan/randomVectorGenerator.cpp compiled into oracle/_ref/libref.so",
        "normal_stddev1_first64": [float(x) for x in oracle.ref_random_vectors(64)],
        "normal_stddev_1_over_32_first64": sd32,
        "signs_first64": [int(x) for x in oracle.ref_random_signs(64)],
    }


def goldstandard_small() -> None:
    rng = np.random.default_rng(20240611)
    data = rng.standard_normal((300, 12)).astype(np.float32)
    queries = rng.standard_normal((6, 12)).astype(np.float32)
    K = 20
    with tempfile.TemporaryDirectory() as td:
        dp, qp, tp = (os.path.join(td, n) for n in ("data.txt", "queries.txt", "truth.txt"))
        oracle.write_text_dataset(dp, data)
        oracle.write_text_dataset(qp, queries)
        subprocess.run([oracle.REF_GOLDSTANDARD, str(K), dp, qp, tp], check=True, capture_output=True)
        ids, dist = oracle.read_truth_file(tp)
        # The text files round-trip float32 exactly (repr of the float64 value).
        np.savez_compressed(os.path.join(HERE, "goldstandard_small.npz"), data=data, queries=queries,
                            truth_ids=ids, truth_dist=dist, K=K)


def validation_small() -> dict:
    rng = np.random.default_rng(7)
    Q, KT, k, report_k = 4, 40, 64, 32
    truth_ids = np.stack([rng.permutation(500)[:KT] for _ in range(Q)]).astype(np.int32)
    truth_d = np.sort(rng.random((Q, KT)).astype(np.float32), axis=1)
    res_ids = np.empty((Q, k), dtype=np.int32)
    for q in range(Q):
        keep = truth_ids[q, : 20 + 3 * q]
        other = np.setdiff1d(np.arange(500), truth_ids[q])[: k - keep.size]
        res_ids[q] = np.concatenate([keep, other])[rng.permutation(k)]
    res_ids[1, 5] = res_ids[1, 4]  # a repeated ID inside the first reportK
    res_d = rng.random((Q, k)).astype(np.float32)
    pts = np.empty((Q, k), dtype=[("ID", np.int32), ("distance", np.float32)])
    pts["ID"], pts["distance"] = res_ids, res_d
    with tempfile.TemporaryDirectory() as td:
        tp = os.path.join(td, "truth.txt")
        with open(tp, "w") as f:
            f.write(f"{Q}\n{KT}\n")
            for q in range(Q):
                f.write(f"{q}:\n")
                for j in range(KT):
                    f.write(f"{truth_ids[q, j]} {repr(float(truth_d[q, j]))}\n")
        recall, ratio = oracle.ref_run_validation(tp, pts, Q, k, report_k)
    return {
        "source": "LinearScan/validation.cpp runValidation compiled into oracle/_ref/libref.so",
        "report_k": report_k, "k": k,
        "truth_ids": truth_ids.tolist(), "truth_dist": [[float(v) for v in r] for r in truth_d],
        "result_ids": res_ids.tolist(), "result_dist": [[float(v) for v in r] for r in res_d],
        "recall": recall, "distance_ratio": ratio,
    }


def main() -> None:
    oracle.build(ref=True)
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref is not built (the reference sources are needed)")
    with open(os.path.join(HERE, "rng_streams.json"), "w") as f:
        json.dump(rng_streams(), f, indent=1)
    goldstandard_small()
    with open(os.path.join(HERE, "validation_small.json"), "w") as f:
        json.dump(validation_small(), f, indent=1)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
