# SPDX-License-Identifier: Apache-2.0
"""Pins the CPU oracle (oracle/benes_oracle.c) before anything is checked
against it.  Each check is chosen so that a plausible slip in the restatement
(a dropped term, a swapped side, an off-by-one in a packet prefix) fails:

* the reference's three golden blobs (tests/golden/*.benes, data files of the
  reference's own test suite, tests/test_benes.cpp:598-637);
* the reference's exact payload-size table (tests/test_benes.cpp:116-125) and
  its "placement does not change the size" rule (:301-306);
* closed forms: identity has no crossed switch and every rank is c!-1; a
  reversal of odd size has an all-zero payload; of even size, input switches
  0 and output switches 1 at every level with ranks 0;
* the factoradic code is a bijection onto [0, c!) that decodes back
  (brute force for c <= 7) and has the published widths;
* every blob round-trips through the restated decoder (forward and inverse),
  including all permutations of sizes 1..7;
* byte equality with the reference library compiled from its own sources
  (oracle/_ref, when built) over many sizes, widths and formats, and with the
  reference-produced digests in tests/golden/fixtures.json.
"""
from __future__ import annotations

import hashlib
import itertools
import json
import math
import os

import numpy as np
import pytest

import support as S

B = S.benes_hip


def read_golden(name):
    with open(os.path.join(S.GOLDEN, name), "rb") as f:
        return f.read()


def header_len(blob):
    return 20 + blob[18] + blob[19]


def test_golden_identity_and_reverse():
    ident = np.arange(32, dtype=np.uint32)
    rev = np.arange(33, dtype=np.uint32)[::-1].copy()
    assert S.oracle_compress(ident) == read_golden("identity-32.benes")
    assert S.oracle_compress(rev) == read_golden("reverse-33.benes")


def test_golden_random47_via_decode():
    # std::shuffle is implementation-defined, so the permutation is recovered
    # by decoding the golden blob, then re-encoded.
    blob = read_golden("random-47.benes")
    perm = S.oracle_decode(blob).astype(np.uint32)
    assert sorted(perm.tolist()) == list(range(47))
    assert S.oracle_compress(perm) == blob


@pytest.mark.parametrize("size,payload", [(31, 15), (32, 15), (33, 16), (63, 37), (64, 38),
                                          (65, 39), (96, 64), (127, 91), (128, 91), (129, 92)])
def test_exact_payload_sizes(size, payload):
    perm = np.random.default_rng(size).permutation(size).astype(np.uint32)
    blob = S.oracle_compress(perm)
    assert len(blob) - header_len(blob) == payload


def test_placement_keeps_size():
    for n in (3, 17, 65, 257):
        perm = np.random.default_rng(n).permutation(n).astype(np.uint32)
        for g in (2, 3, 4, 5):
            sizes = {len(S.oracle_compress(perm, g, pl)) for pl in (0, 1, 2)}
            assert len(sizes) == 1


def test_factoradic_widths_and_bijection():
    lib = S.oracle()
    for c in range(0, 33):
        f = math.factorial(c) - 1
        assert lib.orc_factoradic_bits(c) == f.bit_length()
    for c in range(1, 8):
        seen = set()
        for p in itertools.permutations(range(c)):
            arr = np.array(p, np.uint32)
            out = np.zeros(2, np.uint64)
            lib.orc_encode_factoradic(arr.ctypes.data, c, out.ctypes.data)
            r = int(out[0]) | (int(out[1]) << 64)
            assert 0 <= r < math.factorial(c)
            seen.add(r)
            back = np.zeros(32, np.uint32)
            lib.orc_decode_factoradic(int(out[0]), int(out[1]), c, back.ctypes.data)
            assert tuple(back[:c]) == p
        assert len(seen) == math.factorial(c)
    # identity -> c!-1, reverse -> 0 at c = 32 (128-bit range)
    for c in (20, 32):
        out = np.zeros(2, np.uint64)
        lib.orc_encode_factoradic(np.arange(c, dtype=np.uint32).ctypes.data, c, out.ctypes.data)
        assert int(out[0]) | (int(out[1]) << 64) == math.factorial(c) - 1
        lib.orc_encode_factoradic(np.arange(c, dtype=np.uint32)[::-1].copy().ctypes.data, c,
                                  out.ctypes.data)
        assert int(out[0]) == 0 and int(out[1]) == 0


def _sections(blob, n, g=5, placement=0, ic=(), oc=()):
    L = S.oracle_layout(n, g, placement, ic, oc)
    return L, blob[L.input_offset:L.output_offset], blob[L.output_offset:L.middle_offset], \
        blob[L.middle_offset:L.end_offset]


def _rank_stream_bits(mid, n, g):
    # ranks are concatenated in leaf order; walk the leaf sizes left to right
    sizes = []

    def leaves(c):
        if c <= (1 << g):
            sizes.append(c)
        else:
            leaves((c + 1) // 2)
            leaves(c // 2)

    leaves(n)
    bits = int.from_bytes(mid, "little")
    out, pos = [], 0
    for c in sizes:
        w = (math.factorial(c) - 1).bit_length()
        out.append((c, (bits >> pos) & ((1 << w) - 1)))
        pos += w
    return out


@pytest.mark.parametrize("n", [1, 5, 32, 33, 64, 65, 100, 127, 255, 1000, 4097])
def test_closed_form_identity(n):
    blob = S.oracle_compress(np.arange(n, dtype=np.uint32))
    L, ins, outs, mid = _sections(blob, n)
    assert not any(ins) and not any(outs)
    for c, r in _rank_stream_bits(mid, n, 5):
        assert r == math.factorial(c) - 1


def _bit(blob, byte_off, bit):
    return (blob[byte_off + bit // 8] >> (bit % 8)) & 1


@pytest.mark.parametrize("n", [33, 64, 65, 99, 100, 255, 256, 1001, 4096, 4097])
def test_closed_form_reverse(n):
    # A reversal routes into two reversals.  At a node of odd size every input
    # and output switch is 0; at even size inputs are 0 and outputs 1.  Every
    # rank is 0.  The ones in the output section are exactly the even nodes'
    # output switches.
    blob = S.oracle_compress(np.arange(n, dtype=np.uint32)[::-1].copy())
    L = S.oracle_layout(n)
    checked, even_pairs = 0, 0

    def walk(depth, path, c):
        nonlocal checked, even_pairs
        if c <= 32:
            return
        ia = S.oracle_address(L, 0, depth, path)
        oa = S.oracle_address(L, 1, depth, path)
        for pair in range(c // 2):
            for a, off, want in ((ia, L.input_offset, 0), (oa, L.output_offset, int(c % 2 == 0))):
                packet = pair // a.pair_span
                bit = (a.packet_base + packet * a.full_packet_bits +
                       (a.tail_packet_prefix if packet == a.tail_packet else a.full_packet_prefix) +
                       pair % a.pair_span)
                assert _bit(blob, off, bit) == want
                checked += 1
        if c % 2 == 0:
            even_pairs += c // 2
        walk(depth + 1, path * 2, (c + 1) // 2)
        walk(depth + 1, path * 2 + 1, c // 2)

    walk(0, 0, n)
    _, ins, outs, mid = _sections(blob, n)
    assert checked == 2 * S.oracle().orc_switch_bits(n, 5)
    assert not any(ins)
    assert bin(int.from_bytes(outs, "little")).count("1") == even_pairs
    assert all(r == 0 for _, r in _rank_stream_bits(mid, n, 5))


@pytest.mark.parametrize("n", range(1, 8))
def test_exhaustive_round_trip(n):
    for p in itertools.permutations(range(n)):
        perm = np.array(p, np.uint32)
        blob = S.oracle_compress(perm)
        fwd, inv = S.oracle_decode(blob, inverse=True)
        assert fwd.tolist() == list(p)
        assert inv[perm].tolist() == list(range(n))


def test_round_trip_formats():
    rng = np.random.default_rng(7)
    for n in (3, 7, 15, 17, 31, 33, 63, 65, 127, 129, 257, 1000):
        perm = rng.permutation(n).astype(np.uint32)
        for g in (1, 2, 3, 4, 5):
            d = S.routing_depth(n, g)
            ic = [1] * d
            oc = [d] if d else []
            for pl in (0, 1, 2):
                blob = S.oracle_compress(perm, g, pl, ic, oc)
                assert S.oracle_decode(blob).tolist() == perm.tolist()


def test_errors():
    with pytest.raises(ValueError):
        S.oracle_compress(np.zeros(0, np.uint32))
    with pytest.raises(ValueError):
        S.oracle_compress(np.array([0, 1, 1], np.uint32))
    with pytest.raises(ValueError):
        S.oracle_compress(np.array([0, 1, 3], np.uint64))
    perm = np.random.default_rng(1).permutation(4097).astype(np.uint32)
    with pytest.raises(ValueError):
        S.oracle_compress(perm, ic=(1, 1))
    with pytest.raises(ValueError):
        S.oracle_compress(perm, g=6)
    with pytest.raises(OverflowError):
        S.oracle_layout(2 ** 64 - 1)


@pytest.mark.skipif(not S.ref_available(), reason="oracle/_ref not built (needs reference sources)")
def test_against_reference_library():
    rng = np.random.default_rng(11)
    for n in list(range(1, 70)) + [127, 128, 129, 255, 257, 1023, 1025, 4095, 4097, 65537]:
        for g in (1, 3, 5):
            d = S.routing_depth(n, g)
            splits = [()] + ([tuple([1] * d)] if d else []) + ([(d - 1, 1), (1, d - 1)] if d >= 2 else [])
            perm = rng.permutation(n).astype(np.uint32)
            for pl in (0, 1, 2):
                for ic in splits:
                    oc = splits[-1]
                    a = S.oracle_compress(perm, g, pl, ic, oc)
                    assert a == S.ref_compress(perm, g, pl, ic, oc), (n, g, pl, ic, oc)
                    if n % 7 == 0:
                        assert S.oracle_compress(perm.astype(np.uint64), g, pl, ic, oc) == a


def load_fixtures():
    with open(os.path.join(S.GOLDEN, "fixtures.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("rec", [r for r in load_fixtures() if r["n"] <= (1 << 20) + 1],
                         ids=lambda r: r["name"])
def test_oracle_matches_reference_digests(rec):
    perm = B.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"])
    assert [int(x) for x in perm[:8]] == rec["first8"]
    blob = S.oracle_compress(perm, rec["g"], rec["placement"], rec["ic"], rec["oc"])
    assert len(blob) == rec["size"]
    assert hashlib.sha256(blob).hexdigest() == rec["sha256"]


def load_structured():
    with open(os.path.join(S.GOLDEN, "structured.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("rec", [r for r in load_structured() if r["n"] <= (1 << 17) + 1],
                         ids=lambda r: r["name"])
def test_oracle_matches_reference_structured(rec):
    # Inputs whose level graphs are long chains (support.structured).
    perm = S.structured(rec["kind"], rec["n"])
    blob = S.oracle_compress(perm, rec["g"], rec["placement"], rec["ic"], rec["oc"])
    assert len(blob) == rec["size"]
    assert hashlib.sha256(blob).hexdigest() == rec["sha256"]
