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
