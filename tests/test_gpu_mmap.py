# SPDX-License-Identifier: Apache-2.0
"""Ingest from and egress to memory-mapped files through the host entry point
(benes_hip_compress): the permutation is read from the mapped input file's
pages by the staging copies and DMA'd to HBM chunk by chunk, and the blob is
written from the pinned ring straight into the mapped output file's pages, so
neither side is copied whole on the host.  This covers what the reference's
compress_file (src/benes.cpp:498-553) reads and writes, for raw little-endian
files, with the blobs checked against reference digests and the oracle."""
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


def mapped_compress(tmp_path, perm: np.ndarray, fmt) -> bytes:
    src, dst = tmp_path / "perm.bin", tmp_path / "blob.bin"
    perm.astype(perm.dtype.newbyteorder("<"), copy=False).tofile(src)
    values = np.memmap(src, dtype=perm.dtype, mode="r")
    size = B.encoded_size(perm.size, fmt)
    out = np.memmap(dst, dtype=np.uint8, mode="w+", shape=(size,))
    blob = B.compress_into(values, fmt, out)
    assert blob.size == size
    out.flush()
    del blob, out, values
    return dst.read_bytes()


def test_mapped_files_reference_digest(tmp_path):
    with open(os.path.join(S.GOLDEN, "fixtures.json")) as f:
        rec = next(r for r in json.load(f) if r["name"] == "cfg1-fy-2^20")
    perm = B.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"])
    blob = mapped_compress(tmp_path, perm, S.fmt(rec["g"], rec["placement"], rec["ic"], rec["oc"]))
    assert len(blob) == rec["size"]
    assert hashlib.sha256(blob).hexdigest() == rec["sha256"]


@pytest.mark.parametrize("n,dtype,g,pl", [(53, np.uint64, 5, 2), (300007, np.uint32, 4, 1),
                                          (1 << 23, np.uint64, 5, 0)])
def test_mapped_files_match_oracle(tmp_path, n, dtype, g, pl):
    perm = np.random.default_rng(n).permutation(n).astype(dtype)
    blob = mapped_compress(tmp_path, perm, S.fmt(g, pl))
    if n <= 300007:
        assert blob == S.oracle_compress(perm.astype(np.uint64), g, pl)
    else:
        assert blob == B.compress(perm, S.fmt(g, pl))
