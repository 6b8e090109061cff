# SPDX-License-Identifier: Apache-2.0
"""File ingest and egress (the reference's compress_file, src/benes.cpp:
498-553): raw little-endian uint32/uint64 files are read whole, compressed
through the host path of benes_hip_compress, and the blob written out.  The
blobs equal the oracle's and the reference digests; the reference's errors
come back in its order with its messages, and a failed compress leaves no
output file.  (An earlier build streamed both files through the pinned ring;
one run of it stopped making progress, the cause was not found, and that code
is gone.)"""
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


def write_values(path, perm: np.ndarray) -> None:
    perm.astype(perm.dtype.newbyteorder("<"), copy=False).tofile(path)


@pytest.mark.parametrize("n,width,g,pl", [(1, 4, 5, 0), (47, 8, 5, 2), (1000, 4, 3, 1),
                                          (100003, 8, 4, 2), (300007, 4, 5, 0)])
def test_file_matches_oracle(tmp_path, n, width, g, pl):
    rng = np.random.default_rng(n)
    perm = rng.permutation(n).astype(np.uint32 if width == 4 else np.uint64)
    src, dst = tmp_path / "perm.bin", tmp_path / "blob.bin"
    write_values(src, perm)
    size = B.compress_file(src, dst, width, S.fmt(g, pl))
    blob = dst.read_bytes()
    assert len(blob) == size
    assert blob == S.oracle_compress(perm.astype(np.uint64), g, pl)


def test_file_reference_digest_2e26(tmp_path):
    # 256 MB in, 207 MB out: several ring chunks each way.
    with open(os.path.join(S.GOLDEN, "fixtures.json")) as f:
        rec = next(r for r in json.load(f) if r["name"] == "cfg2-fy-2^26")
    perm = B.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"])
    src, dst = tmp_path / "perm.bin", tmp_path / "blob.bin"
    write_values(src, perm)
    del perm
    size = B.compress_file(src, dst, 4, S.fmt(rec["g"], rec["placement"], rec["ic"], rec["oc"]))
    assert size == rec["size"] == os.path.getsize(dst)
    h = hashlib.sha256()
    with open(dst, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 24), b""):
            h.update(chunk)
    assert h.hexdigest() == rec["sha256"]


def test_file_errors(tmp_path):
    dst = tmp_path / "blob.bin"
    with pytest.raises(RuntimeError, match="cannot open permutation input file"):
        B.compress_file(tmp_path / "missing.bin", dst, 4)
    odd = tmp_path / "odd.bin"
    odd.write_bytes(b"\0" * 6)
    with pytest.raises(ValueError, match="invalid element width or size"):
        B.compress_file(odd, dst, 4)
    empty = tmp_path / "empty.bin"
    empty.write_bytes(b"")
    with pytest.raises(ValueError, match="invalid element width or size"):
        B.compress_file(empty, dst, 4)
    good = tmp_path / "good.bin"
    write_values(good, np.arange(16, dtype=np.uint32))
    with pytest.raises(ValueError, match="invalid element width or size"):
        B.compress_file(good, dst, 3)
    # Not a permutation: the compress error, and no output file.
    dup = tmp_path / "dup.bin"
    write_values(dup, np.array([0, 1, 1, 3], dtype=np.uint32))
    with pytest.raises(ValueError, match="every value from zero"):
        B.compress_file(dup, dst, 4)
    assert not dst.exists()
    # Format errors come after the read, as compress raises them.
    with pytest.raises(ValueError):
        B.compress_file(good, dst, 4, B.FormatOptions(9))
    assert not dst.exists()
    with pytest.raises(RuntimeError, match="failed to write compressed permutation file"):
        B.compress_file(good, tmp_path / "no_such_dir" / "blob.bin", 4)
    # A good call after the failures still works and truncates an old file.
    dst.write_bytes(b"x" * 100000)
    size = B.compress_file(good, dst, 4)
    assert dst.read_bytes() == S.oracle_compress(np.arange(16, dtype=np.uint64)) and size < 100000
