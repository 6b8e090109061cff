# SPDX-License-Identifier: Apache-2.0
"""Byte parity of the gfx950 backend with the oracle (a restatement of the
reference's single-threaded portable encoder) and with digests of blobs the
reference library itself produced (tests/golden/fixtures.json).  Every test
calls through the C ABI (libbenes_hip.so); nothing falls back to the CPU.

Cases follow the reference's own suite (tests/test_benes.cpp): golden blobs,
exhaustive small sizes, boundary sizes, every format option and placement,
uint32 == uint64 bytes, thread/launch determinism, and invalid input.
"""
from __future__ import annotations

import hashlib
import itertools
import json
import os
import re

import numpy as np
import pytest

import support as S

B = S.benes_hip
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    # On a GPU run the backend must be there: fail, never skip.
    assert B.available(), "no gfx950 device or libbenes_hip.so failed to load its kernels"


def read_golden(name):
    with open(os.path.join(S.GOLDEN, name), "rb") as f:
        return f.read()


def test_golden_blobs():
    assert B.compress(np.arange(32, dtype=np.uint32)) == read_golden("identity-32.benes")
    assert B.compress(np.arange(33, dtype=np.uint32)[::-1].copy()) == read_golden("reverse-33.benes")
    blob = read_golden("random-47.benes")
    perm = S.oracle_decode(blob).astype(np.uint32)
    assert B.compress(perm) == blob
    assert B.compress(perm.astype(np.uint64)) == blob


@pytest.mark.parametrize("n", range(1, 7))
def test_exhaustive_small(n):
    for p in itertools.permutations(range(n)):
        perm = np.array(p, np.uint32)
        assert B.compress(perm) == S.oracle_compress(perm), p


BOUNDARY = [9, 15, 17, 30, 31, 33, 34, 47, 62, 63, 65, 66, 95, 96, 97, 126, 127, 129, 130, 255,
            257, 511, 513, 1023, 1025, 4095, 4097]


@pytest.mark.parametrize("n", BOUNDARY)
def test_boundaries(n):
    rng = np.random.default_rng(0xC04FAC70 + n)
    for kind in range(3):
        perm = {0: rng.permutation(n), 1: np.arange(n), 2: np.arange(n)[::-1]}[kind]
        perm = np.ascontiguousarray(perm, dtype=np.uint32)
        assert B.compress(perm) == S.oracle_compress(perm), (n, kind)


@pytest.mark.parametrize("n", [3, 7, 15, 17, 31, 33, 63, 65, 127, 129, 257, 1000, 4097])
def test_format_options(n):
    rng = np.random.default_rng(0xF04A7000 + n)
    perm = rng.permutation(n).astype(np.uint32)
    for g in (1, 2, 3, 4, 5):
        d = S.routing_depth(n, g)
        splits = [()] + ([tuple([1] * d)] if d else []) + ([(d - 1, 1), (1, d - 1), (d,)] if d >= 2 else [])
        sizes = set()
        for pl in (0, 1, 2):
            for ic in splits:
                oc = splits[(len(ic) + pl) % len(splits)]
                got = B.compress(perm, S.fmt(g, pl, ic, oc))
                assert got == S.oracle_compress(perm, g, pl, ic, oc), (n, g, pl, ic, oc)
                if ic == () and oc == ():
                    sizes.add(len(got))
        assert len(sizes) == 1


def test_reference_format_case_4097():
    # tests/test_benes.cpp:318-343: g = 4, input {3, 6}, output {5, 4}, embedded_input
    perm = np.random.default_rng(0xF04A7FFF).permutation(4097).astype(np.uint32)
    args = (4, 1, (3, 6), (5, 4))
    assert B.compress(perm, S.fmt(*args)) == S.oracle_compress(perm, *args)


@pytest.mark.parametrize("n", [1, 2, 3, 31, 32, 33, 63, 65, 127, 129, 4097, 70001])
def test_uint64_same_bytes(n):
    # tests/test_benes.cpp:230-266
    perm = np.random.default_rng(0x64000000 + n).permutation(n).astype(np.uint32)
    d = S.routing_depth(n, 4)
    args = (4, 1 if n % 2 == 0 else 2, tuple([1] * d), (d,) if d else ())
    ref = S.oracle_compress(perm, *args)
    assert B.compress(perm, S.fmt(*args)) == ref
    assert B.compress(perm.astype(np.uint64), S.fmt(*args)) == ref


@pytest.mark.parametrize("n", [16383, 16384, 16385, 32767, 32769, 65537, 100003, 262145])
def test_global_levels(n):
    # Sizes above one workgroup's LDS subtree exercise the HBM union-find,
    # resolve and packet-packing kernels, including odd nodes and tail packets.
    rng = np.random.default_rng(n)
    for g, pl in ((5, 0), (3, 1), (4, 2)):
        d = S.routing_depth(n, g)
        for ic, oc in (((), ()), (tuple([1] * d), (d,)), ((2, d - 3, 1), (d - 1, 1))):
            perm = rng.permutation(n).astype(np.uint32)
            assert B.compress(perm, S.fmt(g, pl, ic, oc)) == S.oracle_compress(perm, g, pl, ic, oc)


@pytest.mark.parametrize("wide", [0, 1], ids=["words32", "words64"])
def test_sixteen_bit_level_arrays(wide):
    # Nodes of at most 2^16 entries keep their level arrays as 16-bit entries
    # inside their own range of the shared 32-bit arrays; the last depth
    # coloured in HBM writes its children that way only where every node of
    # the item's next depth is that small.  Sizes on both sides of 2^16 and
    # 2^17, one at a time and mixed in one batch (one depth then holds HBM
    # nodes of some items and LDS nodes of others), with 32- and 64-bit
    # union-find words and level-0 inverse.
    sizes = [1 << 16, (1 << 16) + 1, (1 << 17) - 1, 1 << 17, (1 << 17) + 1, (1 << 18) + 7]
    rng = np.random.default_rng(0x16B)
    B.set_wide_words(wide)
    try:
        perms, fmts, want = [], [], []
        for i, n in enumerate(sizes):
            width = 8 if i % 2 else 4
            perm = rng.permutation(n).astype(np.uint64 if width == 8 else np.uint32)
            g, pl = (5, 0) if i % 3 else (4, 1)
            ref = S.oracle_compress(perm.astype(np.uint64), g, pl)
            print(f"16-bit levels wide={wide} n={n} width={width}", flush=True)
            assert B.compress(perm, S.fmt(g, pl)) == ref, n
            perms.append(perm)
            fmts.append(S.fmt(g, pl))
            want.append(ref)
        print(f"16-bit levels wide={wide} batch of {len(perms)}", flush=True)
        assert B.compress_batch(perms, fmts) == want
    finally:
        B.set_wide_words(0)


@pytest.mark.parametrize("n", [(1 << 19) - 3, (1 << 20) + 12345])
def test_tall_cluster_groups(n):
    # One cluster group spanning most levels: near its root a packet's share
    # of one node runs to 2^(g-1) bits, and in the packet where node sizes
    # differ (or the group's last, partial packet) each share is written at
    # its node's own address by several threads, word by word.
    rng = np.random.default_rng(0x7A11 + n)
    for g, pl in ((4, 0), (5, 2), (3, 1)):
        d = S.routing_depth(n, g)
        for ic, oc in (((d,), (d,)), ((2, d - 4, 2), (d,)), ((1, d - 1), (d - 2, 2))):
            perm = rng.permutation(n).astype(np.uint32)
            assert B.compress(perm, S.fmt(g, pl, ic, oc)) == S.oracle_compress(perm, g, pl, ic, oc), \
                (n, g, pl, ic, oc)


@pytest.mark.parametrize("shape", B.SUBTREE_SHAPES, ids=lambda s: f"{s[0]}x{s[1]}")
def test_every_subtree_shape(shape):
    # Every subtree shape the tuner may pick gives the same bytes: sizes
    # around each shape's subtree and LDS-level limits, several formats.
    B.set_subtree_shape(*shape)
    try:
        rng = np.random.default_rng(shape[0] * 10 + shape[1])
        for n in (4095, 4096, 4097, 8191, 8193, 16383, 16385, 65537, 100003):
            for g, pl in ((1, 0), (3, 1), (5, 2)):
                d = S.routing_depth(n, g)
                for ic, oc in (((), ()), ((2, d - 3, 1), (d - 1, 1))):
                    perm = rng.permutation(n).astype(np.uint32)
                    want = S.oracle_compress(perm, g, pl, ic, oc)
                    assert B.compress(perm, S.fmt(g, pl, ic, oc)) == want, (n, g, pl, ic, oc)
        rec = next(r for r in load_fixtures() if r["name"] == "cfg1-fy-2^20")
        perm = B.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"])
        assert hashlib.sha256(B.compress(perm)).hexdigest() == rec["sha256"]
    finally:
        B.set_subtree_shape(0, 0)


def _tuner_line(cls: str) -> str:
    return next((l for l in B.autotune_report().splitlines() if l.startswith(cls)), "")


def test_tuner_trials_then_keeps_one_shape():
    # A size class of its own (largest and total 2^21): its first calls cycle
    # through every shape, then the choice is fixed; every call's bytes agree.
    n = (1 << 21) + 4321
    cls = "largest 2^21 total 2^21:"
    perm = B.generate(B.FISHER_YATES, n, 2121)
    want = S.oracle_compress(perm)
    for _ in range(B.AUTOTUNE_TRIAL_CALLS):
        if "->" in _tuner_line(cls):
            break
        assert B.compress(perm) == want
    line = _tuner_line(cls)
    assert "->" in line, line
    times = [float(t) for t in re.findall(r"=([0-9.]+)ms", line)]
    assert len(times) == len(B.SUBTREE_SHAPES) and all(t > 0 for t in times), line
    for _ in range(3):
        assert B.compress(perm) == want
        assert _tuner_line(cls) == line


def test_structured_inputs_large():
    n = 300007
    for kind in (B.SINGLE_CYCLE, B.NEAR_IDENTITY, B.IDENTITY, B.REVERSE):
        perm = B.generate(kind, n, 77, 5000)
        assert B.compress(perm) == S.oracle_compress(perm), kind


def load_fixtures():
    with open(os.path.join(S.GOLDEN, "fixtures.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("rec", load_fixtures(), ids=lambda r: r["name"])
def test_reference_digests(rec):
    perm = B.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"])
    blob = B.compress(perm, S.fmt(rec["g"], rec["placement"], rec["ic"], rec["oc"]))
    assert len(blob) == rec["size"]
    assert hashlib.sha256(blob).hexdigest() == rec["sha256"], rec["name"]


def test_repeatable():
    perm = B.generate(B.FISHER_YATES, 1 << 20, 1)
    first = B.compress(perm)
    for _ in range(3):
        assert B.compress(perm) == first


def test_invalid_input():
    with pytest.raises(ValueError):
        B.compress(np.zeros(0, np.uint32))
    with pytest.raises(ValueError):
        B.compress(np.array([0, 1, 1], np.uint32))
    with pytest.raises(ValueError):
        B.compress(np.array([0, 1, 3], np.uint32))
    with pytest.raises(ValueError):
        B.compress(np.array([0, 1, 1], np.uint64))
    big = np.arange(100000, dtype=np.uint32)
    big[99999] = 5
    with pytest.raises(ValueError):
        B.compress(big)
    perm = np.random.default_rng(1).permutation(4097).astype(np.uint32)
    with pytest.raises(ValueError):
        B.compress(perm, S.fmt(ic=(1, 1)))
    with pytest.raises(ValueError):
        B.compress(perm, S.fmt(g=6))
    # The permutation is checked before the format, as the reference does
    # (src/benes.cpp:361-363): a non-permutation with a bad format reports
    # the permutation, on both entry points.
    with pytest.raises(ValueError, match="every value from zero"):
        B.compress(np.array([0, 1, 1], np.uint32), S.fmt(g=6))
    with pytest.raises(ValueError, match="every value from zero"):
        B.compress(big, S.fmt(ic=(1, 1)))
    with pytest.raises(ValueError, match="middle_group_log2 must be between 1 and 5"):
        B.compress(perm, S.fmt(g=6))
    # the device is still healthy afterwards
    assert B.compress(perm) == S.oracle_compress(perm)


def test_device_entry_point():
    torch = pytest.importorskip("torch")
    n = (1 << 20) + 3
    perm = B.generate(B.FISHER_YATES, n, 5)
    fmt = S.fmt(5, 2)
    host = B.compress(perm, fmt)
    d_perm = torch.from_numpy(perm.view(np.int32)).cuda()
    d_out = torch.zeros(len(host), dtype=torch.uint8, device="cuda")
    size = B.compress_device(d_perm.data_ptr(), n, 4, d_out.data_ptr(), d_out.numel(), fmt,
                             stream=torch.cuda.current_stream().cuda_stream)
    assert size == len(host)
    assert d_out.cpu().numpy().tobytes() == host
