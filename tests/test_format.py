# SPDX-License-Identifier: Apache-2.0
"""Host-side logic of the product, checked on the CPU against the oracle:
the C ABI loads and exports everything include/benes_hip.h declares, layouts
and the closed-form node addressing agree with the oracle's restatement of
SectionAddress (src/internal.hpp:475-718), errors carry the reference's
classes, and the seeded generators are fixed."""
from __future__ import annotations

import json
import os
import random
import re

import numpy as np
import pytest

import support as S

B = S.benes_hip


def test_library_exports_every_declared_symbol():
    lib = B.library()
    with open(os.path.join(S.ROOT, "include", "benes_hip.h")) as f:
        declared = set(re.findall(r"\b(benes_hip_[a-z_]+)\s*\(", f.read()))
    assert declared == set(B.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_header_constants_match_the_module():
    with open(os.path.join(S.ROOT, "include", "benes_hip.h")) as f:
        text = f.read()
    trials = int(re.search(r"#define BENES_HIP_AUTOTUNE_TRIAL_CALLS (\d+)", text).group(1))
    assert trials == B.AUTOTUNE_TRIAL_CALLS
    assert trials % len(B.SUBTREE_SHAPES) == 0
    codes = dict(re.findall(r"#define BENES_HIP_([A-Z_]+) (\d+)", text))
    for name in ("OK", "INVALID_ARGUMENT", "OVERFLOW", "LENGTH", "LOGIC", "RUNTIME",
                 "BUFFER_TOO_SMALL", "PENDING", "OUT_OF_RANGE"):
        assert int(codes[name]) == getattr(B, name), name


def test_layout_matches_oracle():
    for n in list(range(1, 130)) + [255, 4097, 65537, 1 << 20, (1 << 26), (1 << 32) + 12345,
                                    (1 << 40) + 3]:
        for g in (1, 4, 5):
            d = S.routing_depth(n, g)
            for pl in (0, 1, 2):
                for ic in [()] + ([tuple([1] * d)] if d else []):
                    L = S.oracle_layout(n, g, pl, ic, ())
                    lay = B.layout_of(n, S.fmt(g, pl, ic, ()))
                    for k in ("input_offset", "output_offset", "middle_offset", "end_offset",
                              "network_size"):
                        assert lay[k] == getattr(L, k), (n, g, pl, ic, k)
                    assert lay["input_clusters"] == list(L.input_clusters[:L.n_input_clusters])
                    assert lay["output_clusters"] == list(L.output_clusters[:L.n_output_clusters])
                    assert lay["middle_placement"] == L.middle_placement


def test_config_layouts():
    # 2^26 uint32 with input {11,10}, output {10,11}, embedded_input
    lay = B.layout_of(1 << 26, S.fmt(5, 1, (11, 10), (10, 11)))
    assert lay["routing_depth"] == 21
    assert lay["end_offset"] == 207093784
    lay = B.layout_of((1 << 32) + 12345, S.fmt(5, 2))
    assert lay["input_clusters"] == [14, 14] and lay["end_offset"] == 16475259031


KEYS = ("packet_base", "full_packet_bits", "full_packet_prefix", "tail_packet_prefix",
        "pair_span", "tail_packet")


@pytest.mark.parametrize("n", [33, 65, 100, 129, 1000, 4097, 77777, 100003, (1 << 20) + 1,
                               1 << 26, (1 << 32) + 12345])
def test_node_addresses_match_oracle(n):
    rnd = random.Random(n)
    for g in (1, 3, 5):
        d = S.routing_depth(n, g)
        splits = [()] + ([tuple([1] * d)] if d else []) + ([(d - 1, 1), (1, d - 1)] if d >= 2 else [])
        if d >= 4:
            splits.append((2, d - 3, 1))
        for pl in (0, 1, 2):
            ic = rnd.choice(splits)
            oc = rnd.choice(splits)
            L = S.oracle_layout(n, g, pl, ic, oc)
            F = S.fmt(g, pl, ic, oc)
            for depth in range(d + 1):
                paths = range(1 << depth) if depth <= 5 else [rnd.randrange(1 << depth) for _ in range(24)]
                for path in paths:
                    if depth and S.oracle_address(L, 0, depth - 1, path >> 1).count <= (1 << g):
                        continue
                    for side in (0, 1):
                        o = S.oracle_address(L, side, depth, path)
                        a = B.node_address(n, F, side, depth, path)
                        assert a["count"] == o.count
                        if o.count > (1 << g):
                            for k in KEYS:
                                assert a[k] == getattr(o, k), (n, g, pl, ic, oc, side, depth, path, k)
                        elif pl == 0 or pl == side + 1:
                            assert a["rank_bit"] == o.middle_bit


def test_format_errors():
    with pytest.raises(ValueError):
        B.encoded_size(4097, S.fmt(ic=(1, 1)))
    with pytest.raises(ValueError):
        B.encoded_size(4097, S.fmt(g=6))
    with pytest.raises(ValueError):
        B.encoded_size(4097, S.fmt(g=0))
    with pytest.raises(ValueError):
        B.encoded_size(4097, S.fmt(placement=3))
    with pytest.raises(ValueError):
        B.encoded_size(4097, S.fmt(ic=(0, 7)))
    with pytest.raises(OverflowError):
        B.encoded_size(2 ** 64 - 1)
    # a depth-0 permutation forces separate placement, as the reference does
    assert B.layout_of(20, S.fmt(placement=1))["middle_placement"] == 0


def test_generators_are_fixed():
    with open(os.path.join(S.GOLDEN, "fixtures.json")) as f:
        recs = json.load(f)
    for rec in recs:
        if rec["n"] > (1 << 22) + 20000:
            continue
        perm = B.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"])
        assert [int(x) for x in perm[:8]] == rec["first8"]
    for kind in (B.FISHER_YATES, B.SINGLE_CYCLE, B.NEAR_IDENTITY):
        p = B.generate(kind, 1000, 5, 10000)
        assert sorted(p.tolist()) == list(range(1000))
    # Sattolo gives one cycle
    p = B.generate(B.SINGLE_CYCLE, 1000, 9)
    x, steps = 0, 0
    while True:
        x = int(p[x])
        steps += 1
        if x == 0:
            break
    assert steps == 1000
    # near-identity moves at most 2 * floor(rate * n) entries
    p = B.generate(B.NEAR_IDENTITY, 100000, 3, 5000)
    assert (p != np.arange(100000)).sum() <= 2 * 500


def test_keyed_generator_is_a_bijection():
    for n in (1, 2, 3, 5, 17, 1000, 65537, (1 << 20) + 3):
        for width in (4, 8):
            p = B.generate(B.KEYED, n, 77, 0, width)
            assert np.array_equal(np.sort(p), np.arange(n, dtype=p.dtype))
            assert np.array_equal(p, B.generate(B.KEYED, n, 77, 0, width))
    a, b = B.generate(B.KEYED, 5000, 1), B.generate(B.KEYED, 5000, 2)
    assert not np.array_equal(a, b)
    assert (a != np.arange(5000)).sum() > 4900   # no structure left to speak of


def test_uniform_parallel_generator():
    # A permutation for every size, the same for both widths, and (with several
    # buckets, n >= 2^19) no position favoured for a given value.
    for n in (1, 2, 3, 1000, (1 << 19) + 3, (1 << 20) + 1):
        p = B.generate(B.UNIFORM_PARALLEL, n, 11, 0, 8)
        assert np.array_equal(np.sort(p), np.arange(n, dtype=np.uint64)), n
        assert np.array_equal(B.generate(B.UNIFORM_PARALLEL, n, 11, 0, 4).astype(np.uint64), p)
    n = 1 << 20
    quarter = [int(np.flatnonzero(B.generate(B.UNIFORM_PARALLEL, n, s, 0, 4) == 12345)[0]) >> 18
               for s in range(400)]
    assert np.bincount(quarter, minlength=4).min() > 60   # about 100 each
    fixed = [int((B.generate(B.UNIFORM_PARALLEL, n, s, 0, 4) == np.arange(n)).sum())
             for s in range(200)]
    assert 0.6 < np.mean(fixed) < 1.5                      # Poisson(1) fixed points
