# SPDX-License-Identifier: Apache-2.0
"""The ruler colouring of the HBM levels (benes_hip_set_colouring(RULERS):
orbit minima by sampled rulers and pointer jumping, sharing no code with the
union-find) against the oracle and the reference library's digests.  Pinned
here, it is the independent check of the union-find colouring at
2^32 + 12345 entries (tests/test_gpu_large.py), where no CPU encoder runs."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

import support as S

B = S.benes_hip
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def rulers():
    assert B.available(), "no gfx950 device or libbenes_hip.so failed to load its kernels"
    B.set_colouring(B.RULERS)
    yield
    B.set_colouring(B.UNION_FIND)


def compress_profiled(perm, fmt):
    B.set_profiling(True)
    try:
        blob = B.compress(perm, fmt)
        stats = B.kernel_stats()
    finally:
        B.set_profiling(False)
    for kernel in ("k_ruler_walk", "k_ruler_jump", "k_ruler_route"):
        assert stats.get(kernel, (0, 0))[1] > 0, f"{kernel} did not run: {sorted(stats)}"
    assert "k_union" not in stats and "k_hook_first" not in stats
    return blob


@pytest.mark.parametrize("n,g,pl,ic,oc,width", [
    ((1 << 17) + 1, 5, 0, (), (), 4),
    (1 << 20, 5, 1, (), (), 4),
    ((1 << 21) + 12345, 5, 2, (), (), 8),
    ((1 << 22) + 1, 4, 1, (10, 9), (6, 13), 4),
])
def test_rulers_random_match_oracle(n, g, pl, ic, oc, width):
    perm = B.generate(B.FISHER_YATES, n, n ^ 0x8011, 0, width)
    print(f"rulers random n={n} g={g} pl={pl} width={width}", flush=True)
    assert compress_profiled(perm, S.fmt(g, pl, ic, oc)) == S.oracle_compress(perm, g, pl, ic, oc)


@pytest.mark.parametrize("kind", ["reverse", "rotate1", "rotate3", "rotate_half", "shuffle",
                                  "reverse_halves"])
@pytest.mark.parametrize("n", [(1 << 20) + 1, 1 << 21])
def test_rulers_structured_match_oracle(kind, n):
    perm = S.structured(kind, n)
    print(f"rulers structured {kind} n={n}", flush=True)
    assert compress_profiled(perm.astype(np.uint32), S.fmt(5, 1)) == S.oracle_compress(perm, 5, 1)


@pytest.mark.parametrize("kind,rate", [(B.IDENTITY, 0), (B.NEAR_IDENTITY, 5000),
                                       (B.NEAR_IDENTITY, 200000), (B.SINGLE_CYCLE, 0)])
def test_rulers_short_and_long_cycles(kind, rate):
    # The identity and near-identity inputs are mostly orbits without a ruler.
    n = (1 << 20) + 3
    perm = B.generate(kind, n, 31, rate)
    print(f"rulers kind={kind} rate={rate}", flush=True)
    assert compress_profiled(perm, S.fmt(5, 2)) == S.oracle_compress(perm, 5, 2)


def test_rulers_batch():
    rng = np.random.default_rng(77)
    perms = [rng.permutation(n).astype(np.uint32) for n in (1 << 20, 70001, 1000, (1 << 19) + 9)]
    perms.append(S.structured("reverse", (1 << 18) + 1).astype(np.uint32))
    fmts = [S.fmt(5, 0), S.fmt(3, 1), S.fmt(4, 2), S.fmt(5, 2), S.fmt(5, 1)]
    out = B.compress_batch(perms, fmts)
    for p, f, blob in zip(perms, fmts, out):
        assert blob == S.oracle_compress(p, f.middle_group_log2, f.middle_placement)


def digests():
    with open(os.path.join(S.GOLDEN, "fixtures.json")) as f:
        recs = [r for r in json.load(f) if r["n"] > (1 << 16)]
    with open(os.path.join(S.GOLDEN, "structured.json")) as f:
        recs += [dict(r, structured=True) for r in json.load(f) if r["n"] > (1 << 16)]
    return recs


@pytest.mark.parametrize("rec", digests(), ids=lambda r: r["name"])
def test_rulers_reference_digests(rec):
    if rec.get("structured"):
        perm = S.structured(rec["kind"], rec["n"]).astype(np.uint32 if rec["width"] == 4 else np.uint64)
    else:
        perm = B.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"])
    print(f"rulers digest {rec['name']}", flush=True)
    blob = compress_profiled(perm, S.fmt(rec["g"], rec["placement"], rec["ic"], rec["oc"]))
    assert len(blob) == rec["size"]
    assert hashlib.sha256(blob).hexdigest() == rec["sha256"], rec["name"]
