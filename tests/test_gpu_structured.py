# SPDX-License-Identifier: Apache-2.0
"""Structured permutations whose level graphs are long monotone chains.

For a reverse of odd length, a rotation by an odd amount of even length, a
perfect shuffle of odd length or a reversal of each half of odd length, the
first-hook pass (k_hook_first) hangs nearly every input slot under a
neighbouring slot, and the union pass's Rem walk for the edge that closes the
cycle would then climb the whole chain one dependent load at a time: half the
node, 2^31 loads at the root of a 2^32-entry reverse.  Walks are capped, the
capped edges deferred, the forest flattened by all slots at once (pointer
jumping) and the deferred edges joined on it.  The blobs must equal the
oracle's, and a 2^24 rotation must take milliseconds, not the seconds an
end-to-end walk of 2^23 loads costs.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np
import pytest

import support as S

B = S.benes_hip
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    assert B.available(), "no gfx950 device or libbenes_hip.so failed to load its kernels"


KINDS = ["reverse", "rotate1", "rotate3", "rotate_half", "shuffle", "reverse_halves"]


@pytest.mark.parametrize("n", [(1 << 20) + 1, 1 << 21])
@pytest.mark.parametrize("kind", KINDS)
def test_structured_matches_oracle(kind, n):
    perm = S.structured(kind, n)
    assert np.array_equal(np.sort(perm), np.arange(n, dtype=np.uint64))
    print(f"structured {kind} n={n}", flush=True)
    got = B.compress(perm.astype(np.uint32), S.fmt(5, 1))
    assert got == S.oracle_compress(perm, 5, 1)


def test_structured_wide_values():
    n = (1 << 21) + 3
    perm = S.structured("reverse", n)
    got = B.compress(perm, S.fmt(4, 2))
    assert got == S.oracle_compress(perm, 4, 2)


def test_rotation_2e24_is_fast():
    n = 1 << 24
    perm = S.structured("rotate1", n).astype(np.uint32)
    fmt = S.fmt(5, 0)
    B.compress(perm, fmt)
    t0 = time.perf_counter()
    blob = B.compress(perm, fmt)
    elapsed = time.perf_counter() - t0
    print(f"rotation 2^24: {elapsed * 1e3:.1f} ms", flush=True)
    assert elapsed < 2.0
    idx = np.random.default_rng(24).integers(0, n, 4096, dtype=np.uint64)
    assert np.array_equal(B.lookup(blob, idx), (idx + 1) % n)
    assert np.array_equal(B.lookup(blob, idx, inverse=True), (idx + n - 1) % n)


def structured_records():
    with open(os.path.join(S.GOLDEN, "structured.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("rec", structured_records(), ids=lambda r: r["name"])
def test_structured_reference_digests(rec):
    # Digests of the reference library's own blobs at 2^26 scale.
    perm = S.structured(rec["kind"], rec["n"]).astype(np.uint32 if rec["width"] == 4 else np.uint64)
    print(f"structured digest {rec['name']}", flush=True)
    blob = B.compress(perm, S.fmt(rec["g"], rec["placement"], rec["ic"], rec["oc"]))
    assert len(blob) == rec["size"]
    assert hashlib.sha256(blob).hexdigest() == rec["sha256"]


def test_structured_items_in_one_batch():
    # Chain-structured and random items side by side: walk caps, deferred
    # edges and the chain count are per depth over the whole batch.
    perms = [S.structured("rotate1", 1 << 20), S.structured("reverse", (1 << 20) + 1),
             np.random.default_rng(5).permutation(1 << 20).astype(np.uint64),
             S.structured("shuffle", (1 << 19) + 1), S.structured("rotate1", (1 << 21) + 1)]
    fmts = [S.fmt(5, 0), S.fmt(5, 1), S.fmt(4, 2), S.fmt(5, 2), S.fmt(3, 1)]
    out = B.compress_batch([p.astype(np.uint32) for p in perms], fmts)
    for p, f, blob in zip(perms, fmts, out):
        assert blob == S.oracle_compress(p, f.middle_group_log2, f.middle_placement)
