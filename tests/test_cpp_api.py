# SPDX-License-Identifier: Apache-2.0
"""The C++ layer (benes-perm_amd/libbenes.so, include/benes/benes.hpp) as a
user of the reference's API calls it: tests/cpp/test_benes_api links the
library and calls benes::compress (reference include/benes/benes.hpp:47-61,
dispatch src/benes.cpp:422-458).  Without a GPU it checks the status-code to
exception mapping, backend names and the refusal of backends this library
does not serve; on the GPU, the reference's input and format errors in its
order, and blobs against digests of the reference library's own output."""
from __future__ import annotations

import hashlib
import json
import os
import subprocess

import pytest

import support as S

EXE = os.path.join(S.ROOT, "tests", "cpp", "test_benes_api")


def run(*args):
    assert os.path.exists(EXE), "tests/cpp/test_benes_api is not built (make -C benes-perm_amd)"
    proc = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=120)
    fails = [l for l in proc.stdout.splitlines() if l.startswith("FAIL")]
    assert proc.returncode == 0 and not fails, proc.stdout + proc.stderr
    return proc.stdout


def test_status_codes_and_backends():
    out = run("errors")
    assert "ok status 1 -> invalid_argument" in out and "ok backend_name hip" in out


@pytest.mark.gpu
def test_reference_errors_in_order_on_device():
    out = run("errors-device")
    assert "ok duplicate value before a bad format" in out


def fixtures():
    with open(os.path.join(S.GOLDEN, "fixtures.json")) as f:
        names = {"cfg1-fy-2^20", "u64-fy-1060921", "fmt-100003-3-1", "fmt-77777-5-2",
                 "g4-2-3001"}
        return [r for r in json.load(f) if r["name"] in names]


@pytest.mark.gpu
@pytest.mark.parametrize("rec", fixtures(), ids=lambda r: r["name"])
@pytest.mark.parametrize("backend", ["automatic", "hip"])
def test_compress_matches_reference_digest(tmp_path, rec, backend):
    out = tmp_path / "blob.bin"
    lst = lambda xs: ",".join(map(str, xs)) if xs else "-"
    run("compress", out, rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"], rec["g"],
        rec["placement"], lst(rec["ic"]), lst(rec["oc"]), backend)
    blob = out.read_bytes()
    assert len(blob) == rec["size"]
    assert hashlib.sha256(blob).hexdigest() == rec["sha256"], rec["name"]


@pytest.mark.gpu
def test_compress_golden_identity(tmp_path):
    out = tmp_path / "blob.bin"
    run("compress", out, 3, 32, 0, 0, 4, 5, 0, "-", "-", "automatic")   # kind 3: identity
    with open(os.path.join(S.GOLDEN, "identity-32.benes"), "rb") as f:
        assert out.read_bytes() == f.read()


def run_file(src, dst, width=4, g=5, placement=0):
    out = run("file", src, dst, width, g, placement)
    return out.strip().splitlines()[-1]


def write_values(path, perm):
    perm.astype(perm.dtype.newbyteorder("<"), copy=False).tofile(path)


def test_compress_file_errors_before_the_device(tmp_path):
    # The reference's file errors and their order (src/benes.cpp:498-553): the
    # input file, its size, then compress; no output file is created when any
    # of them fails.
    import numpy as np

    dst = tmp_path / "blob.bin"
    assert run_file(tmp_path / "missing.bin", dst) == \
        "threw runtime_error: cannot open permutation input file"
    odd = tmp_path / "odd.bin"
    odd.write_bytes(b"\0" * 6)
    assert run_file(odd, dst) == \
        "threw invalid_argument: permutation input has an invalid element width or size"
    empty = tmp_path / "empty.bin"
    empty.write_bytes(b"")
    assert run_file(empty, dst).startswith("threw invalid_argument")
    good = tmp_path / "good.bin"
    write_values(good, np.arange(16, dtype=np.uint32))
    assert run_file(good, dst, width=3).startswith("threw invalid_argument")
    assert not dst.exists()


@pytest.mark.gpu
@pytest.mark.parametrize("n,width,g,pl", [(1, 4, 5, 0), (47, 8, 5, 2), (1000, 4, 3, 1),
                                          (100003, 8, 4, 2), (300007, 4, 5, 0)])
def test_compress_file_matches_oracle(tmp_path, n, width, g, pl):
    import numpy as np

    perm = np.random.default_rng(n).permutation(n).astype(np.uint32 if width == 4 else np.uint64)
    src, dst = tmp_path / "perm.bin", tmp_path / "blob.bin"
    write_values(src, perm)
    assert run_file(src, dst, width, g, pl) == "ok file"
    assert dst.read_bytes() == S.oracle_compress(perm.astype(np.uint64), g, pl)


@pytest.mark.gpu
def test_compress_file_errors_on_device(tmp_path):
    import numpy as np

    dst = tmp_path / "blob.bin"
    dup = tmp_path / "dup.bin"
    write_values(dup, np.array([0, 1, 1, 3], dtype=np.uint32))
    assert run_file(dup, dst) == ("threw invalid_argument: input must contain every value from "
                                  "zero through size minus one")
    good = tmp_path / "good.bin"
    write_values(good, np.arange(16, dtype=np.uint32))
    assert run_file(good, dst, g=9) == \
        "threw invalid_argument: middle_group_log2 must be between 1 and 5"
    assert not dst.exists()
    assert run_file(good, tmp_path / "no_such_dir" / "blob.bin") == \
        "threw runtime_error: failed to write compressed permutation file"
    dst.write_bytes(b"x" * 100000)   # an old, longer file is truncated
    assert run_file(good, dst) == "ok file"
    assert dst.read_bytes() == S.oracle_compress(np.arange(16, dtype=np.uint64))


@pytest.mark.gpu
def test_compress_file_reference_digest(tmp_path):
    with open(os.path.join(S.GOLDEN, "fixtures.json")) as f:
        rec = next(r for r in json.load(f) if r["name"] == "u64-fy-1060921")
    perm = S.benes_hip.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"])
    src, dst = tmp_path / "perm.bin", tmp_path / "blob.bin"
    write_values(src, perm)
    assert run_file(src, dst, 8, rec["g"], rec["placement"]) == "ok file"
    assert hashlib.sha256(dst.read_bytes()).hexdigest() == rec["sha256"]
