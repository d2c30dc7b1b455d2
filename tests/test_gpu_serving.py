"""The request loop (lsh_server_*): continuous batching of requests submitted
from several threads, each request's results equal to searching its queries
alone; memory admission."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _clustered(rng, n, d, clusters=200):
    centres = rng.standard_normal((clusters, d)).astype(np.float32)
    x = centres[rng.integers(0, clusters, n)] + 0.3 * rng.standard_normal((n, d)).astype(np.float32)
    return (x / np.linalg.norm(x, axis=1, keepdims=True)).astype(np.float32)


def test_server_results_same_alone_and_in_the_mix(lsh):
    rng = np.random.default_rng(900)
    N, d = 100000, 64
    data = _clustered(rng, N, d)
    reqs = []
    for i in range(48):
        n = int(np.ceil(np.exp(rng.uniform(0, np.log(300)))))
        reqs.append((_clustered(rng, n, d), int(rng.choice([1, 10, 32, 100]))))
    out = [None] * len(reqs)
    with lsh.Index(data, implementation=3, sketchDim=16, candidates=256) as ix:
        with lsh.Server(ix, max_batch_queries=1024) as sv:
            def client(j0):
                tickets = [(j, sv.submit(reqs[j][0], reqs[j][1])) for j in range(j0, len(reqs), 4)]
                for j, t in tickets:
                    out[j] = sv.wait(t)
            th = [threading.Thread(target=client, args=(j,)) for j in range(4)]
            for t in th:
                t.start()
            for t in th:
                t.join()
            st = sv.stats()
        assert st["queries"] == sum(len(q) for q, _ in reqs)
        assert st["batches"] < len(reqs)  # requests were batched together
        for (q, k), (res, lat) in zip(reqs, out):
            assert res.shape == (len(q), lsh.calculateK(k)) and lat > 0
            np.testing.assert_array_equal(res, ix.search(q, k))


def test_server_memory_admission(lsh):
    rng = np.random.default_rng(901)
    N, d = 50000, 32
    data = _clustered(rng, N, d)
    with lsh.Index(data, implementation=3, sketchDim=8, candidates=128) as ix:
        # a budget that admits only small batches
        with lsh.Server(ix, max_batch_queries=4096, max_query_size=64, memory_budget_bytes=6 << 20) as sv:
            assert sv.stats()["max_batch"] < 4096
            with pytest.raises(lsh.LshError):
                sv.submit(_clustered(rng, sv.stats()["max_batch"] + 1, d), 32)
            qs = [_clustered(rng, 60, d) for _ in range(10)]
            tickets = [sv.submit(q, 32) for q in qs]
            res = [sv.wait(t)[0] for t in tickets]
            assert sv.stats()["batches"] >= 2
        for q, r in zip(qs, res):
            np.testing.assert_array_equal(r, ix.search(q, 32))
