# SPDX-License-Identifier: Apache-2.0
"""The 64-bit union-find words and the 64-bit level-0 inverse, forced at every
HBM level (benes_hip_set_wide_words(1)), as the reference's Metal backend
forces its 64-bit parents with BENES_METAL_FORCE_PARENT64
(src/metal/benes_metal.mm:615-633, BENCHMARKS.md:163-173).  By default these
instantiations (k_hook_first / k_union / k_flatten / k_route_inputs with
64-bit words, k_scatter_inverse / k_check_inverse / k_route_outputs /
k_level_lds over a 64-bit inverse) only run where a node has 2^31 input slots
or more, i.e. at the root of the 2^32 + 12345 workload.  Here they are checked
byte for byte against the oracle on random, chain-structured and batched
inputs, and against digests of the reference library's own blobs, and a
profile shows that the 64-bit kernels are the ones that ran."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

import support as S

B = S.benes_hip
pytestmark = pytest.mark.gpu

WIDE_KERNELS = ("k_scatter_inverse_w64", "k_check_inverse_w64", "k_hook_first_w64",
                "k_union_w64", "k_flatten_w64", "k_route_inputs_w64", "k_route_outputs_w64")


@pytest.fixture(scope="module", autouse=True)
def wide():
    assert B.available(), "no gfx950 device or libbenes_hip.so failed to load its kernels"
    B.set_wide_words(1)
    yield
    B.set_wide_words(0)


def compress_profiled(perm, fmt):
    B.set_profiling(True)
    try:
        blob = B.compress(perm, fmt)
        stats = B.kernel_stats()
    finally:
        B.set_profiling(False)
    return blob, stats


def assert_wide_ran(stats):
    missing = [k for k in WIDE_KERNELS if stats.get(k, (0, 0))[1] == 0]
    assert not missing, f"64-bit kernels did not run: {missing}; ran {sorted(stats)}"
    narrow = [k for k in ("k_union", "k_hook_first", "k_route_inputs", "k_scatter_inverse")
              if k in stats]
    assert not narrow, f"32-bit instantiations ran in forced-wide mode: {narrow}"


@pytest.mark.parametrize("n,g,pl,ic,oc,width", [
    (1 << 20, 5, 0, (), (), 4),
    ((1 << 21) + 12345, 5, 2, (), (), 8),
    ((1 << 22) + 1, 4, 1, (10, 9), (6, 13), 4),
    (1 << 23, 5, 2, (), (), 8),
])
def test_wide_random_matches_oracle(n, g, pl, ic, oc, width):
    perm = B.generate(B.FISHER_YATES, n, n ^ 0x64, 0, width)
    print(f"wide random n={n} g={g} pl={pl} width={width}", flush=True)
    got, stats = compress_profiled(perm, S.fmt(g, pl, ic, oc))
    assert_wide_ran(stats)
    assert got == S.oracle_compress(perm, g, pl, ic, oc)


@pytest.mark.parametrize("kind", ["reverse", "rotate1", "rotate_half", "shuffle", "reverse_halves"])
def test_wide_structured_matches_oracle(kind):
    # Chains of neighbouring slots: deferred walks and the flatten with
    # 64-bit words, coherent walks at the shallow depths.
    n = (1 << 21) + 1
    perm = S.structured(kind, n)
    print(f"wide structured {kind} n={n}", flush=True)
    got, stats = compress_profiled(perm.astype(np.uint32), S.fmt(5, 1))
    assert_wide_ran(stats)
    assert got == S.oracle_compress(perm, 5, 1)


def test_wide_mixed_batch():
    rng = np.random.default_rng(64)
    perms = [rng.permutation(n).astype(np.uint64 if w == 8 else np.uint32)
             for n, w in ((1 << 20, 4), ((1 << 19) + 7, 8), (70001, 4), (33, 8))]
    perms.append(S.structured("rotate1", (1 << 20) + 1).astype(np.uint32))
    fmts = [S.fmt(5, 0), S.fmt(4, 2), S.fmt(3, 1), S.fmt(5, 2), S.fmt(5, 0)]
    out = B.compress_batch(perms, fmts)
    for p, f, blob in zip(perms, fmts, out):
        assert blob == S.oracle_compress(p.astype(np.uint64), f.middle_group_log2,
                                         f.middle_placement)


def fixtures():
    with open(os.path.join(S.GOLDEN, "fixtures.json")) as f:
        return [r for r in json.load(f) if r["n"] > (1 << 16)]


@pytest.mark.parametrize("rec", fixtures(), ids=lambda r: r["name"])
def test_wide_reference_digests(rec):
    perm = B.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"])
    print(f"wide digest {rec['name']}", flush=True)
    blob, stats = compress_profiled(perm, S.fmt(rec["g"], rec["placement"], rec["ic"], rec["oc"]))
    assert_wide_ran(stats)
    assert len(blob) == rec["size"]
    assert hashlib.sha256(blob).hexdigest() == rec["sha256"], rec["name"]
