# SPDX-License-Identifier: Apache-2.0
"""Test-side bindings: the oracle (oracle/liboracle.so, a C restatement of the
reference encoder) and, when it has been built, the reference library itself
(oracle/_ref/libbenes_ref.so, compiled from the reference sources).

Only tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg use these.
"""
from __future__ import annotations

import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "benes-perm_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

import benes_hip  # noqa: E402

ORACLE_PATH = os.path.join(ROOT, "oracle", "liboracle.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libbenes_ref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")

u64 = ctypes.c_uint64


class OrcLayout(ctypes.Structure):
    _fields_ = [
        ("size", u64), ("network_size", u64),
        ("middle_group_log2", ctypes.c_uint8), ("middle_placement", ctypes.c_uint8),
        ("n_input_clusters", ctypes.c_uint8), ("n_output_clusters", ctypes.c_uint8),
        ("input_clusters", ctypes.c_uint8 * 64), ("output_clusters", ctypes.c_uint8 * 64),
        ("input_offset", u64), ("output_offset", u64), ("middle_offset", u64),
        ("end_offset", u64),
    ]


class OrcAddress(ctypes.Structure):
    _fields_ = [(n, u64) for n in (
        "count", "packet_base", "full_packet_bits", "full_packet_prefix",
        "tail_packet_prefix", "pair_span", "tail_packet", "middle_bit",
        "group_switch_bits")] + [
        ("relative_level", ctypes.c_uint32), ("group_depth", ctypes.c_uint32),
        ("is_leaf", ctypes.c_uint32)]


_oracle = None
_ref = None


def oracle() -> ctypes.CDLL:
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError("oracle/liboracle.so is not built; run `make -C oracle`")
        lib = ctypes.CDLL(ORACLE_PATH)
        vp = ctypes.c_void_p
        lib.orc_compress.argtypes = [vp, u64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                     vp, ctypes.c_uint32, vp, ctypes.c_uint32, vp, u64,
                                     ctypes.POINTER(u64), ctypes.c_char_p, u64]
        lib.orc_make_layout.argtypes = [u64, ctypes.c_uint32, ctypes.c_uint32, vp,
                                        ctypes.c_uint32, vp, ctypes.c_uint32,
                                        ctypes.POINTER(OrcLayout), ctypes.c_char_p, u64]
        lib.orc_lookup.argtypes = [vp, u64, u64, ctypes.c_int, ctypes.POINTER(u64),
                                   ctypes.c_char_p, u64]
        lib.orc_decode_all.argtypes = [vp, u64, vp, vp, ctypes.c_char_p, u64]
        lib.orc_lookup_many.argtypes = [vp, u64, vp, u64, ctypes.c_int, vp, ctypes.c_char_p, u64]
        lib.orc_node_address.argtypes = [ctypes.POINTER(OrcLayout), ctypes.c_int,
                                         ctypes.c_uint32, u64, ctypes.POINTER(OrcAddress)]
        lib.orc_factoradic_bits.argtypes = [ctypes.c_uint32]
        lib.orc_factoradic_bits.restype = ctypes.c_uint32
        lib.orc_encode_factoradic.argtypes = [vp, ctypes.c_uint32, vp]
        lib.orc_decode_factoradic.argtypes = [u64, u64, ctypes.c_uint32, vp]
        for name in ("orc_switch_bits", "orc_middle_bits"):
            getattr(lib, name).argtypes = [u64, ctypes.c_uint32]
            getattr(lib, name).restype = u64
        lib.orc_packet_switch_bits.argtypes = [u64, ctypes.c_uint32, ctypes.c_uint32]
        lib.orc_packet_switch_bits.restype = u64
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref() -> ctypes.CDLL:
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_PATH)
        vp = ctypes.c_void_p
        lib.ref_compress.argtypes = [vp, u64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                     vp, ctypes.c_uint32, vp, ctypes.c_uint32, u64,
                                     ctypes.POINTER(u64), ctypes.c_char_p, u64]
        lib.ref_fetch.argtypes = [vp]
        lib.ref_lookup.argtypes = [vp, u64, u64, ctypes.c_int, ctypes.POINTER(u64)]
        _ref = lib
    return _ref


_EXC = {1: ValueError, 2: OverflowError, 3: MemoryError, 4: AssertionError, 5: RuntimeError,
        7: IndexError}


def _clusters(c):
    c = list(c or ())
    return (ctypes.c_uint8 * max(1, len(c)))(*c), len(c)


def oracle_compress(perm: np.ndarray, g=5, placement=0, ic=(), oc=()) -> bytes:
    perm = np.ascontiguousarray(perm)
    ica, ni = _clusters(ic)
    oca, no = _clusters(oc)
    err = ctypes.create_string_buffer(256)
    size = u64()
    lib = oracle()
    args = (perm.ctypes.data_as(ctypes.c_void_p), perm.size, perm.dtype.itemsize, g, placement,
            ctypes.cast(ica, ctypes.c_void_p), ni, ctypes.cast(oca, ctypes.c_void_p), no)
    rc = lib.orc_compress(*args, None, 0, ctypes.byref(size), err, 256)
    if rc:
        raise _EXC.get(rc, RuntimeError)(err.value.decode())
    out = np.zeros(size.value, np.uint8)
    rc = lib.orc_compress(*args, out.ctypes.data_as(ctypes.c_void_p), out.size,
                          ctypes.byref(size), err, 256)
    if rc:
        raise _EXC.get(rc, RuntimeError)(err.value.decode())
    return out.tobytes()


def ref_compress(perm: np.ndarray, g=5, placement=0, ic=(), oc=(), threads=1) -> bytes:
    perm = np.ascontiguousarray(perm)
    ica, ni = _clusters(ic)
    oca, no = _clusters(oc)
    err = ctypes.create_string_buffer(256)
    size = u64()
    lib = ref()
    rc = lib.ref_compress(perm.ctypes.data_as(ctypes.c_void_p), perm.size, perm.dtype.itemsize,
                          g, placement, ctypes.cast(ica, ctypes.c_void_p), ni,
                          ctypes.cast(oca, ctypes.c_void_p), no, threads, ctypes.byref(size),
                          err, 256)
    if rc:
        raise _EXC.get(rc, RuntimeError)(err.value.decode())
    out = np.zeros(size.value, np.uint8)
    lib.ref_fetch(out.ctypes.data_as(ctypes.c_void_p))
    return out.tobytes()


def oracle_decode(blob: bytes, inverse: bool = False) -> np.ndarray:
    """Forward images of every input (and inverse images when asked)."""
    buf = np.frombuffer(blob, np.uint8)
    n = int.from_bytes(blob[8:16], "little")
    fwd = np.zeros(n, np.uint64)
    inv = np.zeros(n, np.uint64) if inverse else None
    err = ctypes.create_string_buffer(256)
    rc = oracle().orc_decode_all(buf.ctypes.data_as(ctypes.c_void_p), buf.size,
                                 fwd.ctypes.data_as(ctypes.c_void_p),
                                 inv.ctypes.data_as(ctypes.c_void_p) if inverse else None,
                                 err, 256)
    if rc:
        raise _EXC.get(rc, RuntimeError)(err.value.decode())
    return (fwd, inv) if inverse else fwd


def oracle_lookup(blob: np.ndarray, values, inverse: bool = False) -> np.ndarray:
    """Images of `values` under the blob's permutation (or its inverse),
    decoded by the oracle; `blob` is a uint8 array holding the whole blob."""
    vals = np.ascontiguousarray(values, np.uint64)
    res = np.zeros(vals.size, np.uint64)
    err = ctypes.create_string_buffer(256)
    rc = oracle().orc_lookup_many(blob.ctypes.data_as(ctypes.c_void_p), blob.size,
                                  vals.ctypes.data_as(ctypes.c_void_p), vals.size,
                                  1 if inverse else 0, res.ctypes.data_as(ctypes.c_void_p),
                                  err, 256)
    if rc:
        raise _EXC.get(rc, RuntimeError)(err.value.decode())
    return res


def oracle_layout(n, g=5, placement=0, ic=(), oc=()) -> OrcLayout:
    ica, ni = _clusters(ic)
    oca, no = _clusters(oc)
    out = OrcLayout()
    err = ctypes.create_string_buffer(256)
    rc = oracle().orc_make_layout(n, g, placement, ctypes.cast(ica, ctypes.c_void_p), ni,
                                  ctypes.cast(oca, ctypes.c_void_p), no, ctypes.byref(out),
                                  err, 256)
    if rc:
        raise _EXC.get(rc, RuntimeError)(err.value.decode())
    return out


def oracle_address(layout: OrcLayout, side: int, depth: int, path: int) -> OrcAddress:
    out = OrcAddress()
    rc = oracle().orc_node_address(ctypes.byref(layout), side, depth, path, ctypes.byref(out))
    if rc:
        raise _EXC.get(rc, RuntimeError)("node address")
    return out


def structured(kind: str, n: int) -> np.ndarray:
    """Structured permutations of 0..n-1 (uint64) whose level graphs are long
    monotone chains for the union-find: a reverse, rotations by 1, 3 and
    n/2 + 1, a perfect shuffle (evens then odds) and a reversal of each half."""
    i = np.arange(n, dtype=np.uint64)
    if kind == "reverse":
        return i[::-1].copy()
    if kind == "rotate1":
        return (i + 1) % n
    if kind == "rotate3":
        return (i + 3) % n
    if kind == "rotate_half":
        return (i + n // 2 + 1) % n
    if kind == "shuffle":
        h = (n + 1) // 2
        return np.where(i < h, 2 * i, 2 * (i - h) + 1).astype(np.uint64)
    if kind == "reverse_halves":
        return np.concatenate([i[: n // 2][::-1], i[n // 2:][::-1]])
    raise ValueError(f"unknown structured permutation {kind!r}")


def fmt(g=5, placement=0, ic=(), oc=()):
    return benes_hip.FormatOptions(g, tuple(ic), tuple(oc), placement)


def routing_depth(n: int, g: int) -> int:
    d = 0
    while n > (1 << g):
        n = (n + 1) // 2
        d += 1
    return d
