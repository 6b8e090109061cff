# SPDX-License-Identifier: Apache-2.0
"""The request loop: requests of mixed sizes, widths and formats submitted
from several threads while earlier ones run; every result must be the bytes
of that request compressed alone, the loop must actually batch and overlap,
and requests beyond the HBM budget are refused."""
from __future__ import annotations

import threading

import numpy as np
import pytest

import support as S

B = S.benes_hip
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    assert B.available(), "no gfx950 device or libbenes_hip.so failed to load its kernels"


def request_set(seed, count, max_log2):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(2 ** rng.uniform(8, max_log2))
        width = 8 if rng.random() < 0.2 else 4
        u = rng.random()
        kind = B.FISHER_YATES if u < 0.6 else (B.SINGLE_CYCLE if u < 0.8 else B.NEAR_IDENTITY)
        perm = B.generate(kind, n, 9000 + i, 5000, width)
        g = int(rng.integers(4, 6))
        d = S.routing_depth(n, g)
        if d >= 2 and rng.random() < 0.5:
            cut = int(rng.integers(1, d))
            ic, oc = (cut, d - cut), (d - cut, cut)
        else:
            ic, oc = (), ()
        reqs.append((perm, S.fmt(g, int(rng.integers(0, 3)), ic, oc)))
    return reqs


def test_results_same_alone_and_in_the_mix():
    reqs = request_set(1, 160, 18)
    results = {}
    with B.Server(budget_bytes=2 << 30, max_batch=64, lanes=2) as srv:
        ids = [None] * len(reqs)

        def producer(k):
            for i in range(k, len(reqs), 4):
                ids[i] = srv.submit(*reqs[i])

        threads = [threading.Thread(target=producer, args=(k,)) for k in range(4)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        for i, rid in enumerate(ids):
            results[i] = srv.wait(rid)
        stats = srv.stats()
    for i, (perm, fmt) in enumerate(reqs):
        blob, latency = results[i]
        assert not isinstance(blob, Exception), blob
        assert latency > 0
        assert blob.tobytes() == B.compress(perm, fmt), i
    assert stats["completed"] == len(reqs)
    assert stats["batches"] < len(reqs)            # requests were batched
    assert stats["peak_bytes"] <= stats["budget_bytes"]


def test_small_results_match_oracle():
    reqs = request_set(2, 40, 13)
    with B.Server(budget_bytes=1 << 30, max_batch=16, lanes=3) as srv:
        ids = [srv.submit(p, f) for p, f in reqs]
        out = [srv.wait(i)[0].tobytes() for i in ids]
    for (perm, fmt), blob in zip(reqs, out):
        assert blob == S.oracle_compress(perm, fmt.middle_group_log2, fmt.middle_placement,
                                         fmt.input_clusters, fmt.output_clusters)


def test_budget_and_bad_requests():
    with B.Server(budget_bytes=64 << 20, max_batch=8, lanes=2) as srv:
        with pytest.raises(MemoryError):
            srv.submit(B.generate(B.FISHER_YATES, 1 << 21, 1))     # beyond one lane's share
        dup = np.arange(5000, dtype=np.uint32)
        dup[3] = 4
        bad = srv.submit(dup)
        good = srv.submit(B.generate(B.FISHER_YATES, 5000, 2))
        blob, _ = srv.wait(bad)
        assert isinstance(blob, ValueError)
        blob, _ = srv.wait(good)
        assert blob.tobytes() == S.oracle_compress(B.generate(B.FISHER_YATES, 5000, 2))
        with pytest.raises(ValueError):
            srv.submit(np.zeros(0, np.uint32))


def test_large_requests_overlap_under_a_64gb_budget():
    # Requests up to 2^24 entries, uint64 among them, under the served
    # workload's 64 GB budget: lanes run batches at the same time, and every
    # blob is the request's own plain compress.
    rng = np.random.default_rng(24)
    reqs = []
    for i in range(24):
        n = int(2 ** rng.uniform(16, 24)) if i % 6 else (1 << 24) - int(rng.integers(0, 257))
        width = 8 if i % 4 == 1 or i == 12 else 4
        kind = (B.FISHER_YATES, B.SINGLE_CYCLE, B.NEAR_IDENTITY)[i % 3]
        reqs.append((B.generate(kind, n, 24000 + i, 5000, width), S.fmt(5, i % 3)))
    assert any(p.size >= (1 << 24) - 256 and p.dtype == np.uint64 for p, _ in reqs)
    with B.Server(budget_bytes=64 << 30, max_batch=64, lanes=4) as srv:
        ids = [srv.submit(p, f) for p, f in reqs]
        results = [srv.wait(i) for i in ids]
        stats = srv.stats()
    for (perm, fmt), (blob, _) in zip(reqs, results):
        assert not isinstance(blob, Exception), blob
        assert blob.tobytes() == B.compress(perm, fmt)
    assert stats["max_concurrent_batches"] > 1
    assert stats["peak_bytes"] <= stats["budget_bytes"]
