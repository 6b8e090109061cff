# SPDX-License-Identifier: Apache-2.0
"""GPU lookup (the reference's forward_lookup / inverse_lookup,
src/internal.hpp:1036-1139): blobs written by the oracle (the restated
portable encoder) decode on the device to the permutation and its inverse,
as the oracle's own decoder reads them, for every placement, cluster split,
leaf size and odd shape; malformed headers fail with the reference's
messages and out-of-range values with IndexError (std::out_of_range)."""
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
def device():
    assert B.available(), "no gfx950 device or libbenes_hip.so failed to load its kernels"


# (n, middle_group_log2, placement, cluster split): 0 default, 1 every level
# its own cluster, 2 three uneven clusters in, two out.
CASES = [(1, 5, 0, 0), (2, 5, 0, 0), (33, 5, 0, 0), (47, 5, 2, 0), (64, 1, 1, 1),
         (97, 3, 2, 2), (1000, 4, 1, 1), (4097, 5, 0, 2), (65537, 2, 2, 2), (100003, 5, 1, 0),
         (300007, 4, 2, 2)]


def fix(case):
    n, g, pl, split = case
    d = S.routing_depth(n, g)
    ic, oc = (), ()
    if split == 1 and d >= 1:
        ic, oc = tuple([1] * d), (d,)
    elif split == 2 and d >= 4:
        ic, oc = (2, d - 3, 1), (d - 1, 1)
    return n, g, pl, ic, oc


@pytest.mark.parametrize("case", [fix(c) for c in CASES], ids=lambda c: f"{c[0]}-g{c[1]}-p{c[2]}")
def test_lookup_matches_permutation(case):
    n, g, pl, ic, oc = case
    rng = np.random.default_rng(n)
    perm = rng.permutation(n).astype(np.uint32)
    blob = S.oracle_compress(perm, g, pl, ic, oc)
    values = np.arange(n, dtype=np.uint64)
    fwd = B.lookup(blob, values)
    assert np.array_equal(fwd, perm.astype(np.uint64))
    inv = B.lookup(blob, values, inverse=True)
    assert np.array_equal(inv[perm], values)
    assert np.array_equal(fwd, S.oracle_decode(blob))


def test_lookup_structured_and_u64():
    for kind in (B.IDENTITY, B.REVERSE, B.SINGLE_CYCLE, B.NEAR_IDENTITY):
        perm = B.generate(kind, 70001, 5, 5000, 8)
        blob = B.compress(perm, S.fmt(4, 2))
        assert np.array_equal(B.lookup(blob, np.arange(70001, dtype=np.uint64)), perm)


def test_lookup_errors():
    perm = np.random.default_rng(3).permutation(1000).astype(np.uint32)
    blob = S.oracle_compress(perm)
    with pytest.raises(IndexError, match="out of range"):
        B.lookup(blob, [0, 1000])
    bad = bytearray(blob)
    bad[0] ^= 1
    with pytest.raises(ValueError, match="invalid magic"):
        B.lookup(bytes(bad), [0])
    with pytest.raises(ValueError, match="truncated"):
        B.lookup(blob[:10], [0])
    with pytest.raises(ValueError, match="section bounds"):
        B.lookup(blob + b"\0", [0])
    bad = bytearray(blob)
    bad[16] = 3
    with pytest.raises(ValueError, match="format parameters"):
        B.lookup(bytes(bad), [0])


def test_device_round_trip_2e26():
    # The whole 2^26 workload blob decodes back to its input on the device.
    torch = pytest.importorskip("torch")
    with open(os.path.join(S.GOLDEN, "fixtures.json")) as f:
        rec = next(r for r in json.load(f) if r["name"] == "cfg2-fy-2^26")
    perm = B.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"])
    fmt = S.fmt(rec["g"], rec["placement"], rec["ic"], rec["oc"])
    d_perm = torch.from_numpy(perm.view(np.int32)).cuda()
    size = B.encoded_size(rec["n"], fmt)
    d_blob = torch.empty((size + 7) // 8 * 8, dtype=torch.uint8, device="cuda")
    B.compress_device(d_perm.data_ptr(), rec["n"], 4, d_blob.data_ptr(), d_blob.numel(), fmt)
    torch.cuda.synchronize()
    assert hashlib.sha256(d_blob[:size].cpu().numpy().tobytes()).hexdigest() == rec["sha256"]
    d_dec = torch.empty(rec["n"], dtype=torch.int32, device="cuda")
    B.lookup_device(d_blob.data_ptr(), size, None, rec["n"], d_dec.data_ptr(), 4)
    torch.cuda.synchronize()
    assert torch.equal(d_dec, d_perm)
    B.lookup_device(d_blob.data_ptr(), size, None, rec["n"], d_dec.data_ptr(), 4, inverse=True)
    torch.cuda.synchronize()
    assert torch.equal(d_perm[d_dec.long()], torch.arange(rec["n"], dtype=torch.int32,
                                                          device="cuda"))
