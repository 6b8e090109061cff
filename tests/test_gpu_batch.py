# SPDX-License-Identifier: Apache-2.0
"""Batched compression: many permutations of uneven sizes, widths, formats
and structures in one call.  Every blob must equal the oracle's bytes for
that item alone (and hence the plain compress of it), and a bad item must
fail on its own without disturbing the others."""
from __future__ import annotations

import numpy as np
import pytest

import support as S

B = S.benes_hip
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def device():
    assert B.available(), "no gfx950 device or libbenes_hip.so failed to load its kernels"


def random_format(rng, n):
    g = int(rng.integers(1, 6))
    d = S.routing_depth(n, g)
    choice = int(rng.integers(0, 4))
    if d == 0 or choice == 0:
        ic, oc = (), ()
    elif choice == 1:
        ic, oc = tuple([1] * d), (d,)
    else:
        cut = int(rng.integers(1, d + 1))
        ic = (cut, d - cut) if cut < d else (d,)
        cut = int(rng.integers(1, d + 1))
        oc = (d - cut, cut) if cut < d else (d,)
    return g, int(rng.integers(0, 3)), ic, oc


def make_items(seed, count, max_log2):
    rng = np.random.default_rng(seed)
    perms, fmts, args = [], [], []
    for i in range(count):
        n = int(2 ** rng.uniform(0, max_log2)) + int(rng.integers(0, 3))
        kind = int(rng.integers(0, 3))
        width = 8 if rng.random() < 0.3 else 4
        p = B.generate(kind, n, 1000 + i, 5000, width)
        g, pl, ic, oc = random_format(rng, n)
        perms.append(p)
        fmts.append(S.fmt(g, pl, ic, oc))
        args.append((g, pl, ic, oc))
    return perms, fmts, args


def test_mixed_batch_matches_oracle():
    perms, fmts, args = make_items(1, 120, 17)
    out = B.compress_batch(perms, fmts)
    for p, a, blob in zip(perms, args, out):
        assert not isinstance(blob, Exception), blob
        assert blob == S.oracle_compress(p, *a), (p.size, a)


def test_batch_equals_plain_calls():
    perms, fmts, _ = make_items(2, 24, 20)
    out = B.compress_batch(perms, fmts)
    for p, f, blob in zip(perms, fmts, out):
        assert blob == B.compress(p, f)


def test_bad_items_fail_alone():
    good = B.generate(B.FISHER_YATES, 5000, 3)
    dup = np.arange(3000, dtype=np.uint32)
    dup[5] = 7
    rng_bad = np.arange(100, dtype=np.uint64)
    rng_bad[10] = 1000
    perms = [good, dup, good, rng_bad, good]
    fmts = [None, None, S.fmt(g=7), None, S.fmt(4, 2)]
    out = B.compress_batch(perms, fmts)
    assert out[0] == S.oracle_compress(good)
    assert isinstance(out[1], ValueError)
    assert isinstance(out[2], ValueError)
    assert isinstance(out[3], ValueError)
    assert out[4] == S.oracle_compress(good, 4, 2)


def test_device_batch_packed_output():
    torch = pytest.importorskip("torch")
    perms, fmts, args = make_items(4, 64, 16)
    # Pack all inputs into one device tensor at 8-byte aligned offsets.
    offs, total = [], 0
    for p in perms:
        offs.append(total)
        total += (p.nbytes + 7) // 8 * 8
    host = np.zeros(total, np.uint8)
    for p, o in zip(perms, offs):
        host[o:o + p.nbytes] = p.view(np.uint8)
    d_in = torch.from_numpy(host).cuda()
    ns = [p.size for p in perms]
    widths = [p.dtype.itemsize for p in perms]
    out_offs, sizes, need = B.batch_layout(ns, widths, fmts)
    d_out = torch.zeros(need, dtype=torch.uint8, device="cuda")
    ptrs = [d_in.data_ptr() + o for o in offs]
    got_offs, got_sizes, errs, secs = B.compress_batch_device(
        ptrs, ns, widths, fmts, d_out.data_ptr(), need, torch.cuda.current_stream().cuda_stream)
    assert got_offs == out_offs and got_sizes == sizes
    assert all(e is None for e in errs)
    res = d_out.cpu().numpy()
    for p, a, o, sz in zip(perms, args, got_offs, got_sizes):
        assert res[o:o + sz].tobytes() == S.oracle_compress(p, *a)


def test_batch_workload_shape():
    # The batched workload's shape at reduced count: log-uniform sizes in
    # [2^10, 2^20], 30% shifted off powers of two, 60/20/20 structures,
    # middle_group_log2 4 or 5, separate placement.
    rng = np.random.default_rng(77)
    perms, fmts = [], []
    for i in range(48):
        n = int(2 ** rng.uniform(10, 20))
        if rng.random() < 0.3:
            n += int(rng.integers(1, 257))
        u = rng.random()
        kind = B.FISHER_YATES if u < 0.6 else (B.SINGLE_CYCLE if u < 0.8 else B.NEAR_IDENTITY)
        perms.append(B.generate(kind, n, 5000 + i, 5000))
        fmts.append(S.fmt(int(rng.integers(4, 6)), 0))
    out = B.compress_batch(perms, fmts)
    for p, f, blob in zip(perms, fmts, out):
        assert blob == S.oracle_compress(p, f.middle_group_log2, 0)


@pytest.mark.parametrize("plain", [-1, 0], ids=["plain-store", "cas"])
def test_lane_groups_match_oracle(plain):
    # Enough items of 2^17..2^18 entries that their HBM depths take lane
    # groups (each small node on one XCD, several items' nodes interleaved
    # per XCD, partial last groups, lanes of unequal chunk counts), with the
    # union pass in either form; one bad item among them fails alone.
    rng = np.random.default_rng(0x1A4E)
    perms, fmts, args = [], [], []
    for i in range(72):
        n = int(2 ** rng.uniform(17, 18)) + int(rng.integers(0, 4))
        kind = int(rng.integers(0, 3))
        width = 8 if rng.random() < 0.25 else 4
        perms.append(B.generate(kind, n, 7000 + i, 5000, width))
        g, pl, ic, oc = random_format(rng, n)
        fmts.append(S.fmt(g, pl, ic, oc))
        args.append((g, pl, ic, oc))
    bad = np.arange((1 << 17) + 5, dtype=np.uint32)
    bad[77] = 78
    perms.insert(30, bad)
    fmts.insert(30, None)
    args.insert(30, None)
    B.set_plain_union(plain)
    B.set_profiling(True)
    try:
        out = B.compress_batch(perms, fmts)
        stats = B.kernel_stats()
    finally:
        B.set_profiling(False)
        B.set_plain_union(-1)
    assert stats.get("lane_group_depths", (0, 0))[1] >= 2, stats
    assert (stats.get("k_union_plain", (0, 0))[1] > 0) == (plain == -1), stats
    for p, a, blob in zip(perms, args, out):
        if a is None:
            assert isinstance(blob, ValueError)
            continue
        assert not isinstance(blob, Exception), blob
        assert blob == S.oracle_compress(p, *a), (p.size, a)
