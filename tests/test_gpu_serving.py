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


def _unit(rng, n, d):
    x = rng.standard_normal((n, d)).astype(np.float32)
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


def _serve_all(sv, reqs, clients=4):
    out = [None] * len(reqs)

    def client(j0):
        tickets = [(j, sv.submit(reqs[j][0], reqs[j][1])) for j in range(j0, len(reqs), clients)]
        for j, t in tickets:
            out[j] = sv.wait(t)

    th = [threading.Thread(target=client, args=(j,)) for j in range(clients)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    return out


def test_server_captured_batches(lsh):
    # The SimHash index with pivot cells at k <= 64: the loop enqueues its
    # batches through the index's captured graphs, several on the device at
    # once.  One request holds a query whose pass overflows (1,500 copies of
    # its point at distance 0): the select flags it and the loop redoes it.
    # Every request gets what searching its queries alone gives.
    rng = np.random.default_rng(902)
    N, d = 120000, 48
    centres = rng.standard_normal((400, d)).astype(np.float32)
    data = centres[rng.integers(0, 400, N)] + 0.3 * rng.standard_normal((N, d)).astype(np.float32)
    data /= np.linalg.norm(data, axis=1, keepdims=True)
    copies = rng.choice(N, 1500, replace=False)
    data[copies] = data[copies[0]]
    reqs = []
    for i in range(64):
        n = int(np.ceil(np.exp(rng.uniform(0, np.log(400)))))
        q = centres[rng.integers(0, 400, n)] + 0.3 * rng.standard_normal((n, d)).astype(np.float32)
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        reqs.append((q.astype(np.float32), int(rng.choice([1, 10, 32, 64]))))
    reqs[7][0][0] = data[copies[0]]
    with lsh.Index(data, implementation=3, sketchDim=16, distanceFunc=1, candidates=256) as ix:
        with lsh.Server(ix, max_batch_queries=1024) as sv:
            out = _serve_all(sv, reqs)
            st = sv.stats()
        bs = ix.batch_stats()
        assert bs["graph_batches"] == st["batches"] and bs["redone_queries"] >= 1, (bs, st)
        assert st["queries"] == sum(len(q) for q, _ in reqs) and st["batches"] < len(reqs)
        for (q, k), (res, lat) in zip(reqs, out):
            assert res.shape == (len(q), lsh.calculateK(k)) and lat > 0
            np.testing.assert_array_equal(res, ix.search(q, k))
    assert (out[7][0]["ID"][0] == np.sort(copies)[: lsh.calculateK(reqs[7][1])]).all()


def test_server_captured_batches_dense_cells_redone(lsh):
    # Unclustered data: the first captured batch finds the cells' second
    # pass too dense; the loop redoes it by an ordinary search and the index
    # stops using the captured form.
    rng = np.random.default_rng(903)
    data = _unit(rng, 100000, 32)
    reqs = [(_unit(rng, int(n), 32), 32) for n in rng.integers(1, 120, 24)]
    with lsh.Index(data, implementation=3, sketchDim=8, distanceFunc=1, candidates=256) as ix:
        with lsh.Server(ix, max_batch_queries=1024) as sv:
            out = _serve_all(sv, reqs, clients=2)
        bs = ix.batch_stats()
        assert bs["graph_batches"] >= 1 and bs["redone_queries"] > 0, bs
        for (q, k), (res, _) in zip(reqs, out):
            np.testing.assert_array_equal(res, ix.search(q, k))
