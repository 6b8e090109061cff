# SPDX-License-Identifier: Apache-2.0
"""First-hook binning (k_hook_bin, k_hook_scatter): the random first-hook
stores of very large nodes are staged in bins of union-find words, each then
written by one XCD.  By default only levels of 2^31 entries or more bin; the
test setting bins every node whose words exceed 4 MB, so the binned path is
checked here against the oracle and the reference digests at sizes the oracle
finishes in seconds, on random and chain-structured inputs, alone and in a
batch with small items."""
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
def binning():
    assert B.available(), "no gfx950 device or libbenes_hip.so failed to load its kernels"
    B.set_hook_binning(1)
    yield
    B.set_hook_binning(-1)


@pytest.fixture(autouse=True)
def binned_kernels_ran(request):
    # Every compress test here must have gone through the binned kernels, not
    # silently through k_hook_first.
    B.set_profiling(True)
    yield
    stats = B.kernel_stats()
    B.set_profiling(False)
    if request.node.name.startswith(("test_binning_modes", "test_binned_inverse_bad")):
        return
    for kernel in ("k_hook_bin", "k_hook_scatter", "k_inverse_bin", "k_inverse_scatter"):
        assert stats.get(kernel, (0, 0))[1] > 0, f"{kernel} did not run: {sorted(stats)}"


@pytest.mark.parametrize("n,g,pl,width", [((1 << 21) + 5, 5, 0, 4), ((1 << 22) + 3, 4, 2, 4),
                                          ((1 << 23) + 1, 5, 1, 8), (1 << 23, 3, 2, 4)])
def test_binned_random_matches_oracle(n, g, pl, width):
    perm = np.random.default_rng(n + g).permutation(n).astype(np.uint64)
    print(f"binned random n={n} g={g} pl={pl} width={width}", flush=True)
    got = B.compress(perm.astype(np.uint32 if width == 4 else np.uint64), S.fmt(g, pl))
    assert got == S.oracle_compress(perm, g, pl)


@pytest.mark.parametrize("kind", ["reverse", "rotate1", "shuffle", "reverse_halves"])
def test_binned_structured_matches_oracle(kind):
    # Chains of neighbouring slots: the first hooks of the binned path must
    # still feed the adjacency count that switches the walks to coherent ones.
    n = (1 << 22) + 1
    perm = S.structured(kind, n)
    print(f"binned structured {kind} n={n}", flush=True)
    assert B.compress(perm.astype(np.uint32), S.fmt(5, 1)) == S.oracle_compress(perm, 5, 1)


def test_binned_batch_matches_oracle():
    perms = [np.random.default_rng(7).permutation((1 << 22) + 7).astype(np.uint64),
             np.random.default_rng(8).permutation(1000).astype(np.uint64),
             S.structured("rotate3", 1 << 22),
             np.random.default_rng(9).permutation((1 << 21) + 1).astype(np.uint64)]
    fmts = [S.fmt(5, 0), S.fmt(4, 1), S.fmt(5, 2), S.fmt(3, 0)]
    out = B.compress_batch([p.astype(np.uint32) for p in perms], fmts)
    for p, f, blob in zip(perms, fmts, out):
        assert blob == S.oracle_compress(p, f.middle_group_log2, f.middle_placement)


def test_binned_reference_digests_2e26():
    with open(os.path.join(S.GOLDEN, "fixtures.json")) as f:
        recs = [r for r in json.load(f) if r["n"] == 1 << 26]
    with open(os.path.join(S.GOLDEN, "structured.json")) as f:
        structured = [r for r in json.load(f) if r["n"] >= 1 << 26]
    assert recs and structured
    for rec in recs:
        perm = B.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"])
        print(f"binned digest {rec['name']}", flush=True)
        blob = B.compress(perm, S.fmt(rec["g"], rec["placement"], rec["ic"], rec["oc"]))
        assert hashlib.sha256(blob).hexdigest() == rec["sha256"]
    for rec in structured:
        perm = S.structured(rec["kind"], rec["n"]).astype(np.uint32)
        print(f"binned digest {rec['name']}", flush=True)
        blob = B.compress(perm, S.fmt(rec["g"], rec["placement"], rec["ic"], rec["oc"]))
        assert hashlib.sha256(blob).hexdigest() == rec["sha256"]


def test_binning_modes_are_checked():
    with pytest.raises(ValueError):
        B.set_hook_binning(2)


def test_binned_inverse_bad_item_fails_alone():
    # A repeated value overfills a bin of the binned inverse; the bin is shared
    # with the neighbouring items, which must still come out right (the whole
    # inverse is then redone directly and only the bad item is blamed).
    good1 = np.random.default_rng(11).permutation((1 << 20) + 5).astype(np.uint32)
    bad = np.zeros((1 << 20) + 3, np.uint32)              # every value 0
    good2 = np.random.default_rng(12).permutation(1 << 20).astype(np.uint32)
    bad2 = np.arange(1 << 19, dtype=np.uint32)
    bad2[-1] = 0                                          # one repeat, no overflow
    out = B.compress_batch([good1, bad, good2, bad2], [S.fmt(5, 0)] * 4)
    assert out[0] == S.oracle_compress(good1.astype(np.uint64))
    assert isinstance(out[1], ValueError) and "every value from zero" in str(out[1])
    assert out[2] == S.oracle_compress(good2.astype(np.uint64))
    assert isinstance(out[3], ValueError)
