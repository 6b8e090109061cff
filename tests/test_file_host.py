# SPDX-License-Identifier: Apache-2.0
"""benes_hip_compress_file's host side, without a device: the reference's
checks before compression (src/benes.cpp:498-553), in its order and with its
messages, and no output file unless compression succeeded.  On a machine
without a gfx950 device a well-formed input reaches compression, which
reports the backend unavailable, and the output is still not created."""
from __future__ import annotations

import numpy as np
import pytest

import support as S

B = S.benes_hip


def write_values(path, perm: np.ndarray) -> None:
    perm.astype(perm.dtype.newbyteorder("<"), copy=False).tofile(path)


def test_checks_before_compression(tmp_path):
    dst = tmp_path / "blob.bin"
    with pytest.raises(RuntimeError, match="cannot open permutation input file"):
        B.compress_file(tmp_path / "missing.bin", dst, 4)
    odd = tmp_path / "odd.bin"
    odd.write_bytes(b"\0" * 6)
    with pytest.raises(ValueError, match="invalid element width or size"):
        B.compress_file(odd, dst, 4)
    with pytest.raises(ValueError, match="invalid element width or size"):
        B.compress_file(odd, dst, 8)
    empty = tmp_path / "empty.bin"
    empty.write_bytes(b"")
    with pytest.raises(ValueError, match="invalid element width or size"):
        B.compress_file(empty, dst, 4)
    good = tmp_path / "good.bin"
    write_values(good, np.arange(16, dtype=np.uint32))
    with pytest.raises(ValueError, match="invalid element width or size"):
        B.compress_file(good, dst, 3)
    # A directory opens but is not a permutation file.
    with pytest.raises((RuntimeError, ValueError)):
        B.compress_file(tmp_path, dst, 4)
    assert not dst.exists()


def test_format_error_after_the_read(tmp_path):
    good = tmp_path / "good.bin"
    write_values(good, np.arange(40, dtype=np.uint64))
    dst = tmp_path / "blob.bin"
    with pytest.raises(ValueError, match="middle_group_log2"):
        B.compress_file(good, dst, 8, B.FormatOptions(9))
    assert not dst.exists()


@pytest.mark.skipif(B.available(), reason="a device is present: tests/test_gpu_file.py covers it")
def test_no_device_leaves_no_output(tmp_path):
    good = tmp_path / "good.bin"
    write_values(good, np.arange(16, dtype=np.uint32))
    dst = tmp_path / "blob.bin"
    dst.write_bytes(b"old")
    with pytest.raises(RuntimeError, match="unavailable"):
        B.compress_file(good, dst, 4)
    assert dst.read_bytes() == b"old"
