# SPDX-License-Identifier: Apache-2.0
"""ctypes binding of the gfx950 compression backend (include/benes_hip.h).

Mirrors the reference's compression interface: ``compress(permutation,
format)`` returns the compact version-1 blob of a uint32/uint64 permutation,
``FormatOptions`` carries ``middle_group_log2``, the two cluster lists and the
middle placement (include/benes/benes.hpp:32-39 of the reference), and errors
are raised as the Python analogue of the reference's exception class:
ValueError for std::invalid_argument, OverflowError for std::overflow_error,
MemoryError for std::length_error, AssertionError for std::logic_error and
RuntimeError for std::runtime_error.

The shared library must have been built (``make -C benes-perm_amd``); there is
no fallback of any kind.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os
from typing import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbenes_hip.so")

(OK, INVALID_ARGUMENT, OVERFLOW, LENGTH, LOGIC, RUNTIME, BUFFER_TOO_SMALL, PENDING,
 OUT_OF_RANGE) = range(9)

SEPARATE, EMBEDDED_INPUT, EMBEDDED_OUTPUT = 0, 1, 2
FISHER_YATES, SINGLE_CYCLE, NEAR_IDENTITY, IDENTITY, REVERSE, KEYED, UNIFORM_PARALLEL = range(7)

# Every extern "C" symbol declared in include/benes_hip.h.
EXPORTS = (
    "benes_hip_available",
    "benes_hip_compress",
    "benes_hip_compress_device",
    "benes_hip_compress_device_noalloc",
    "benes_hip_compress_file",
    "benes_hip_batch_layout",
    "benes_hip_compress_batch",
    "benes_hip_compress_batch_device",
    "benes_hip_server_create",
    "benes_hip_server_submit",
    "benes_hip_server_wait",
    "benes_hip_server_stats",
    "benes_hip_server_destroy",
    "benes_hip_generate_device",
    "benes_hip_release_workspace",
    "benes_hip_autotune_report",
    "benes_hip_set_subtree_shape",
    "benes_hip_set_hook_binning",
    "benes_hip_set_wide_words",
    "benes_hip_set_plain_union",
    "benes_hip_set_colouring",
    "benes_hip_lookup",
    "benes_hip_lookup_device",
    "benes_hip_encoded_size",
    "benes_hip_last_device_seconds",
    "benes_hip_set_profiling",
    "benes_hip_kernel_stats",
    "benes_hip_layout_of",
    "benes_hip_node_address",
    "benes_hip_address_bit",
    "benes_hip_generate",
)


class _Format(ctypes.Structure):
    _fields_ = [
        ("middle_group_log2", ctypes.c_uint8),
        ("middle_placement", ctypes.c_uint8),
        ("n_input_clusters", ctypes.c_uint8),
        ("n_output_clusters", ctypes.c_uint8),
        ("input_clusters", ctypes.POINTER(ctypes.c_uint8)),
        ("output_clusters", ctypes.POINTER(ctypes.c_uint8)),
    ]


class _Item(ctypes.Structure):
    _fields_ = [
        ("perm", ctypes.c_void_p),
        ("n", ctypes.c_uint64),
        ("width_bytes", ctypes.c_uint32),
        ("status", ctypes.c_int32),
        ("format", _Format),
        ("out", ctypes.c_void_p),
        ("out_capacity", ctypes.c_uint64),
        ("out_offset", ctypes.c_uint64),
        ("out_size", ctypes.c_uint64),
        ("message", ctypes.c_char * 96),
    ]


class _Counters(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "completed", "entries", "batches", "peak_bytes", "max_concurrent_batches", "queued",
        "budget_bytes")]


class _Layout(ctypes.Structure):
    _fields_ = [
        ("size", ctypes.c_uint64),
        ("network_size", ctypes.c_uint64),
        ("input_offset", ctypes.c_uint64),
        ("output_offset", ctypes.c_uint64),
        ("middle_offset", ctypes.c_uint64),
        ("end_offset", ctypes.c_uint64),
        ("switch_bits", ctypes.c_uint64),
        ("middle_bits", ctypes.c_uint64),
        ("routing_depth", ctypes.c_uint32),
        ("middle_group_log2", ctypes.c_uint8),
        ("middle_placement", ctypes.c_uint8),
        ("n_input_clusters", ctypes.c_uint8),
        ("n_output_clusters", ctypes.c_uint8),
        ("input_clusters", ctypes.c_uint8 * 64),
        ("output_clusters", ctypes.c_uint8 * 64),
    ]


class _Address(ctypes.Structure):
    _fields_ = [(name, ctypes.c_uint64) for name in (
        "count", "packet_base", "full_packet_bits", "full_packet_prefix",
        "tail_packet_prefix", "pair_span", "tail_packet", "rank_bit", "offset")]


@dataclasses.dataclass
class FormatOptions:
    """The reference's FormatOptions (include/benes/benes.hpp:32-39)."""
    middle_group_log2: int = 5
    input_clusters: Sequence[int] = ()
    output_clusters: Sequence[int] = ()
    middle_placement: int = SEPARATE


_lib = None


def library() -> ctypes.CDLL:
    """Loads libbenes_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `make -C benes-perm_amd`")
    lib = ctypes.CDLL(LIB_PATH)
    u64, u32, vp, cp, sz = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t
    fp = ctypes.POINTER(_Format)
    lib.benes_hip_available.restype = ctypes.c_int
    lib.benes_hip_compress.argtypes = [vp, u64, u32, fp, vp, u64, ctypes.POINTER(u64),
                                       ctypes.POINTER(ctypes.c_double), cp, sz]
    lib.benes_hip_compress_device.argtypes = [vp, u64, u32, fp, vp, u64, ctypes.POINTER(u64),
                                              vp, cp, sz]
    lib.benes_hip_encoded_size.argtypes = [u64, fp, ctypes.POINTER(u64), cp, sz]
    ip = ctypes.POINTER(_Item)
    lib.benes_hip_batch_layout.argtypes = [ip, u64, ctypes.POINTER(u64)]
    lib.benes_hip_compress_batch.argtypes = [ip, u64, ctypes.POINTER(ctypes.c_double), cp, sz]
    lib.benes_hip_compress_batch_device.argtypes = [ip, u64, vp, u64, vp,
                                                    ctypes.POINTER(ctypes.c_double), cp, sz]
    lib.benes_hip_last_device_seconds.restype = ctypes.c_double
    lib.benes_hip_server_create.argtypes = [u64, u32, u32, cp, sz]
    lib.benes_hip_server_create.restype = vp
    lib.benes_hip_server_submit.argtypes = [vp, vp, u64, u32, fp, vp, u64,
                                            ctypes.POINTER(ctypes.c_int), cp, sz]
    lib.benes_hip_server_submit.restype = u64
    lib.benes_hip_server_wait.argtypes = [vp, u64, ctypes.c_int64, ctypes.POINTER(u64),
                                          ctypes.POINTER(ctypes.c_double), cp, sz]
    lib.benes_hip_server_stats.argtypes = [vp, ctypes.POINTER(_Counters)]
    lib.benes_hip_server_destroy.argtypes = [vp]
    lib.benes_hip_server_destroy.restype = None
    lib.benes_hip_set_profiling.argtypes = [ctypes.c_int]
    lib.benes_hip_set_profiling.restype = None
    lib.benes_hip_kernel_stats.argtypes = [ctypes.c_int, cp, sz, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(u64)]
    lib.benes_hip_layout_of.argtypes = [u64, fp, ctypes.POINTER(_Layout), cp, sz]
    lib.benes_hip_node_address.argtypes = [u64, fp, ctypes.c_int, u32, u64,
                                           ctypes.POINTER(_Address), cp, sz]
    lib.benes_hip_address_bit.argtypes = [ctypes.POINTER(_Address), u64]
    lib.benes_hip_address_bit.restype = u64
    lib.benes_hip_generate.argtypes = [u32, u64, u64, u32, vp, u32]
    lib.benes_hip_generate_device.argtypes = [u32, u64, u64, vp, u32, vp]
    lib.benes_hip_release_workspace.restype = None
    lib.benes_hip_autotune_report.argtypes = [cp, sz]
    lib.benes_hip_set_subtree_shape.argtypes = [u32, u32]
    lib.benes_hip_set_hook_binning.argtypes = [ctypes.c_int]
    lib.benes_hip_set_wide_words.argtypes = [ctypes.c_int]
    lib.benes_hip_set_plain_union.argtypes = [ctypes.c_int]
    lib.benes_hip_set_colouring.argtypes = [ctypes.c_int]
    lib.benes_hip_compress_device_noalloc.argtypes = [vp, u64, u32, fp, vp, u64,
                                                      ctypes.POINTER(u64), vp, cp, sz]
    lib.benes_hip_compress_file.argtypes = [cp, cp, u32, fp, ctypes.POINTER(u64),
                                            ctypes.POINTER(ctypes.c_double), cp, sz]
    lib.benes_hip_lookup.argtypes = [vp, u64, vp, u64, ctypes.c_int, vp, u32, cp, sz]
    lib.benes_hip_lookup_device.argtypes = [vp, u64, vp, u64, ctypes.c_int, vp, u32, vp, cp, sz]
    _lib = lib
    return lib


_ERRORS = {
    INVALID_ARGUMENT: ValueError,
    OVERFLOW: OverflowError,
    LENGTH: MemoryError,
    LOGIC: AssertionError,
    RUNTIME: RuntimeError,
    BUFFER_TOO_SMALL: BufferError,
    OUT_OF_RANGE: IndexError,
}


def _raise(rc: int, err: ctypes.Array) -> None:
    if rc != OK:
        raise _ERRORS.get(rc, RuntimeError)(err.value.decode(errors="replace"))


def _format(fmt: FormatOptions | None):
    fmt = fmt or FormatOptions()
    ic = (ctypes.c_uint8 * max(1, len(fmt.input_clusters)))(*fmt.input_clusters)
    oc = (ctypes.c_uint8 * max(1, len(fmt.output_clusters)))(*fmt.output_clusters)
    f = _Format(fmt.middle_group_log2 & 0xFF, fmt.middle_placement & 0xFF,
                len(fmt.input_clusters), len(fmt.output_clusters),
                ctypes.cast(ic, ctypes.POINTER(ctypes.c_uint8)),
                ctypes.cast(oc, ctypes.POINTER(ctypes.c_uint8)))
    return f, (ic, oc)


def available() -> bool:
    return bool(library().benes_hip_available())


def encoded_size(n: int, fmt: FormatOptions | None = None) -> int:
    f, keep = _format(fmt)
    out = ctypes.c_uint64()
    err = ctypes.create_string_buffer(256)
    _raise(library().benes_hip_encoded_size(n, ctypes.byref(f), ctypes.byref(out), err, 256), err)
    return out.value


def compress_into(perm: np.ndarray, fmt: FormatOptions | None = None,
                  out: np.ndarray | None = None) -> np.ndarray:
    """Compresses a host uint32/uint64 permutation on the GPU into `out`
    (a uint8 array of at least encoded_size bytes, allocated when None);
    returns the blob as a view of it."""
    perm = np.ascontiguousarray(perm)
    if perm.dtype not in (np.uint32, np.uint64):
        raise TypeError("permutation must be uint32 or uint64")
    f, keep = _format(fmt)
    lib = library()
    err = ctypes.create_string_buffer(256)
    size = ctypes.c_uint64()
    secs = ctypes.c_double()
    if out is None:
        # Size query through the compress entry point itself, so that errors
        # come in the reference's order (the permutation before the format).
        rc = lib.benes_hip_compress(perm.ctypes.data_as(ctypes.c_void_p), perm.size,
                                    perm.dtype.itemsize, ctypes.byref(f), None, 0,
                                    ctypes.byref(size), None, err, 256)
        if rc != BUFFER_TOO_SMALL:
            _raise(rc if rc != OK else LOGIC, err)
        out = np.empty(size.value, np.uint8)
    rc = lib.benes_hip_compress(perm.ctypes.data_as(ctypes.c_void_p), perm.size,
                                perm.dtype.itemsize, ctypes.byref(f),
                                out.ctypes.data_as(ctypes.c_void_p), out.size,
                                ctypes.byref(size), ctypes.byref(secs), err, 256)
    _raise(rc, err)
    return out[:size.value]


def compress_file(input_path, output_path, width: int = 4,
                  fmt: FormatOptions | None = None) -> int:
    """The reference's compress_file: raw little-endian values of `width`
    bytes from `input_path`, the blob written to `output_path` (the file is
    read whole and compressed as a host array).  Returns the blob's size."""
    f, keep = _format(fmt)
    size = ctypes.c_uint64()
    err = ctypes.create_string_buffer(256)
    _raise(library().benes_hip_compress_file(os.fsencode(input_path), os.fsencode(output_path),
                                             width, ctypes.byref(f), ctypes.byref(size), None,
                                             err, 256), err)
    del keep
    return size.value


def compress(perm: np.ndarray, fmt: FormatOptions | None = None) -> bytes:
    """Compresses a host uint32/uint64 permutation on the GPU (the blob as
    bytes, as the reference's compress returns it by value)."""
    return compress_into(perm, fmt).tobytes()


def compress_device(d_perm_ptr: int, n: int, width: int, d_out_ptr: int, capacity: int,
                    fmt: FormatOptions | None = None, stream: int | None = None) -> int:
    """Device-resident compress: pointers are device addresses.  Returns the
    blob size."""
    f, keep = _format(fmt)
    err = ctypes.create_string_buffer(256)
    size = ctypes.c_uint64()
    rc = library().benes_hip_compress_device(ctypes.c_void_p(d_perm_ptr), n, width,
                                             ctypes.byref(f), ctypes.c_void_p(d_out_ptr),
                                             capacity, ctypes.byref(size),
                                             ctypes.c_void_p(stream or 0), err, 256)
    _raise(rc, err)
    return size.value


# Limits of benes_hip_compress_device_noalloc (include/benes_hip.h).
NOALLOC_MAX_ENTRIES = 1 << 20
NOALLOC_SCRATCH_BYTES = 36 << 20


def compress_device_noalloc(d_perm_ptr: int, n: int, width: int, d_out_ptr: int, capacity: int,
                            fmt: FormatOptions | None = None, stream: int | None = None) -> int:
    """compress_device without any device allocation: the library's fixed
    per-device scratch holds every buffer (n <= NOALLOC_MAX_ENTRIES; d_out
    8-byte aligned, capacity at least the blob size rounded up to 8).
    Returns the blob size."""
    f, keep = _format(fmt)
    err = ctypes.create_string_buffer(256)
    size = ctypes.c_uint64()
    rc = library().benes_hip_compress_device_noalloc(ctypes.c_void_p(d_perm_ptr), n, width,
                                                     ctypes.byref(f), ctypes.c_void_p(d_out_ptr),
                                                     capacity, ctypes.byref(size),
                                                     ctypes.c_void_p(stream or 0), err, 256)
    del keep
    _raise(rc, err)
    return size.value


def _item_dtype():
    """numpy view of _Item (the same field offsets), for filling and reading
    item tables column by column."""
    f = _Item.format.offset
    fields = {
        "perm": (np.uint64, _Item.perm.offset), "n": (np.uint64, _Item.n.offset),
        "width_bytes": (np.uint32, _Item.width_bytes.offset),
        "status": (np.int32, _Item.status.offset),
        "middle_group_log2": (np.uint8, f + _Format.middle_group_log2.offset),
        "middle_placement": (np.uint8, f + _Format.middle_placement.offset),
        "n_input_clusters": (np.uint8, f + _Format.n_input_clusters.offset),
        "n_output_clusters": (np.uint8, f + _Format.n_output_clusters.offset),
        "input_clusters": (np.uint64, f + _Format.input_clusters.offset),
        "output_clusters": (np.uint64, f + _Format.output_clusters.offset),
        "out": (np.uint64, _Item.out.offset), "out_capacity": (np.uint64, _Item.out_capacity.offset),
        "out_offset": (np.uint64, _Item.out_offset.offset),
        "out_size": (np.uint64, _Item.out_size.offset),
    }
    return np.dtype({"names": list(fields), "formats": [v[0] for v in fields.values()],
                     "offsets": [v[1] for v in fields.values()], "itemsize": ctypes.sizeof(_Item)})


_ITEM_DTYPE = None


def _items(perm_ptrs, ns, widths, fmts):
    """(ctypes _Item array, objects to keep alive, numpy view of the array).
    Formats are converted once per distinct format, and the table is filled
    column by column: a batch of thousands of items spends microseconds here,
    not the tens of milliseconds of per-item ctypes assignments."""
    global _ITEM_DTYPE
    if _ITEM_DTYPE is None:
        _ITEM_DTYPE = _item_dtype()
    count = len(ns)
    rec = np.zeros(count, _ITEM_DTYPE)
    keep = [rec]
    rec["perm"] = [0 if p is None else p for p in perm_ptrs]
    rec["n"] = ns
    rec["width_bytes"] = widths
    cache = {}
    cols = np.zeros((count, 6), np.uint64)
    for i, fmt in enumerate(fmts):
        fmt = fmt or FormatOptions()
        key = (fmt.middle_group_log2, fmt.middle_placement, tuple(fmt.input_clusters),
               tuple(fmt.output_clusters))
        row = cache.get(key)
        if row is None:
            f, k = _format(fmt)
            keep.append(k)
            row = (f.middle_group_log2, f.middle_placement, f.n_input_clusters,
                   f.n_output_clusters, ctypes.cast(f.input_clusters, ctypes.c_void_p).value or 0,
                   ctypes.cast(f.output_clusters, ctypes.c_void_p).value or 0)
            cache[key] = row
        cols[i] = row
    for j, name in enumerate(("middle_group_log2", "middle_placement", "n_input_clusters",
                              "n_output_clusters", "input_clusters", "output_clusters")):
        rec[name] = cols[:, j]
    arr = (_Item * count).from_buffer(rec)
    return arr, keep, rec


def _item_error(item) -> Exception | None:
    if item.status == OK:
        return None
    return _ERRORS.get(item.status, RuntimeError)(item.message.decode(errors="replace"))


def batch_layout(ns, widths, fmts):
    """(out_offsets, out_sizes, total_bytes) of a device batch output."""
    arr, keep, rec = _items([None] * len(ns), ns, widths, fmts)
    total = ctypes.c_uint64()
    library().benes_hip_batch_layout(arr, len(ns), ctypes.byref(total))
    return rec["out_offset"].tolist(), rec["out_size"].tolist(), total.value


def compress_batch(perms, fmts=None):
    """Compresses many host permutations in one call.  Returns, per item, the
    blob bytes or the exception that item raised."""
    fmts = fmts or [None] * len(perms)
    perms = [np.ascontiguousarray(p) for p in perms]
    arr, keep, _ = _items([p.ctypes.data for p in perms], [p.size for p in perms],
                          [p.dtype.itemsize for p in perms], fmts)
    sizes = []
    for it in arr:
        # out_size is filled by the call; allocate from the layout first
        sizes.append(0)
    offs, sizes, _ = batch_layout([p.size for p in perms], [p.dtype.itemsize for p in perms], fmts)
    outs = [np.zeros(max(1, s), np.uint8) for s in sizes]
    for it, o in zip(arr, outs):
        it.out = o.ctypes.data
        it.out_capacity = o.size
    err = ctypes.create_string_buffer(256)
    secs = ctypes.c_double()
    _raise(library().benes_hip_compress_batch(arr, len(perms), ctypes.byref(secs), err, 256), err)
    result = []
    for it, o in zip(arr, outs):
        e = _item_error(it)
        result.append(e if e is not None else o[:it.out_size].tobytes())
    return result


def compress_batch_device(perm_ptrs, ns, widths, fmts, d_out_ptr, capacity, stream=None):
    """Device batch: inputs and the packed output are device memory.  Returns
    (offsets, sizes, errors, device_seconds)."""
    arr, keep, rec = _items(perm_ptrs, ns, widths, fmts)
    err = ctypes.create_string_buffer(256)
    secs = ctypes.c_double()
    _raise(library().benes_hip_compress_batch_device(arr, len(ns), ctypes.c_void_p(d_out_ptr),
                                                     capacity, ctypes.c_void_p(stream or 0),
                                                     ctypes.byref(secs), err, 256), err)
    status = rec["status"]
    errors = [None] * len(ns)
    for i in np.flatnonzero(status != OK):
        errors[i] = _item_error(arr[i])
    return rec["out_offset"].tolist(), rec["out_size"].tolist(), errors, secs.value


class Server:
    """The request loop of include/benes_hip.h (continuous batching over
    `lanes` device streams, admission against an HBM budget)."""

    def __init__(self, budget_bytes: int, max_batch: int = 512, lanes: int = 2):
        err = ctypes.create_string_buffer(256)
        self._h = library().benes_hip_server_create(budget_bytes, max_batch, lanes, err, 256)
        if not self._h:
            raise RuntimeError(err.value.decode())
        self._pending = {}

    def submit(self, perm: np.ndarray, fmt: FormatOptions | None = None) -> int:
        perm = np.ascontiguousarray(perm)
        f, keep = _format(fmt)
        size = encoded_size(perm.size, fmt)
        out = np.empty(size, np.uint8)
        status = ctypes.c_int()
        err = ctypes.create_string_buffer(256)
        rid = library().benes_hip_server_submit(self._h, perm.ctypes.data, perm.size,
                                                perm.dtype.itemsize, ctypes.byref(f),
                                                out.ctypes.data, out.size, ctypes.byref(status),
                                                err, 256)
        if rid == 0:
            _raise(status.value, err)
        self._pending[rid] = (perm, out, keep)
        return rid

    def wait(self, rid: int, timeout_ms: int = -1):
        """(blob as a uint8 array, or the exception of the request; latency
        seconds); raises TimeoutError, keeping the request, when timeout_ms
        runs out."""
        size = ctypes.c_uint64()
        lat = ctypes.c_double()
        err = ctypes.create_string_buffer(256)
        rc = library().benes_hip_server_wait(self._h, rid, timeout_ms, ctypes.byref(size),
                                             ctypes.byref(lat), err, 256)
        if rc == PENDING:
            raise TimeoutError(f"request {rid} still running")
        perm, out, keep = self._pending.pop(rid)
        if rc != OK:
            return _ERRORS.get(rc, RuntimeError)(err.value.decode()), lat.value
        return out[:size.value], lat.value

    def stats(self) -> dict:
        c = _Counters()
        library().benes_hip_server_stats(self._h, ctypes.byref(c))
        return {name: getattr(c, name) for name, _ in _Counters._fields_}

    def close(self):
        if self._h:
            library().benes_hip_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def lookup(blob, values, inverse: bool = False) -> np.ndarray:
    """Images of `values` (host) under the permutation a host blob encodes
    (or under its inverse), decoded on the GPU."""
    buf = np.frombuffer(blob, np.uint8) if isinstance(blob, (bytes, bytearray)) else \
        np.ascontiguousarray(blob, np.uint8)
    vals = np.ascontiguousarray(values, np.uint64)
    out = np.empty(vals.size, np.uint64)
    err = ctypes.create_string_buffer(256)
    _raise(library().benes_hip_lookup(buf.ctypes.data, buf.size, vals.ctypes.data, vals.size,
                                      1 if inverse else 0, out.ctypes.data, 8, err, 256), err)
    return out


def lookup_device(d_blob_ptr: int, blob_size: int, d_values_ptr: int | None, count: int,
                  d_out_ptr: int, width: int = 8, inverse: bool = False,
                  stream: int | None = None) -> None:
    """Device lookups; d_values_ptr None means values 0..count-1."""
    err = ctypes.create_string_buffer(256)
    _raise(library().benes_hip_lookup_device(ctypes.c_void_p(d_blob_ptr), blob_size,
                                             ctypes.c_void_p(d_values_ptr or 0), count,
                                             1 if inverse else 0, ctypes.c_void_p(d_out_ptr),
                                             width, ctypes.c_void_p(stream or 0), err, 256), err)


def last_device_seconds() -> float:
    return library().benes_hip_last_device_seconds()


def set_profiling(enable: bool) -> None:
    """Per-kernel HIP-event timing of later calls (clears the sums)."""
    library().benes_hip_set_profiling(1 if enable else 0)


def kernel_stats() -> dict:
    """{kernel name: (total ms, launches)} since profiling was enabled."""
    lib = library()
    out = {}
    name = ctypes.create_string_buffer(64)
    ms = ctypes.c_double()
    launches = ctypes.c_uint64()
    i = 0
    while lib.benes_hip_kernel_stats(i, name, 64, ctypes.byref(ms), ctypes.byref(launches)):
        out[name.value.decode()] = (ms.value, launches.value)
        i += 1
    return out


def layout_of(n: int, fmt: FormatOptions | None = None) -> dict:
    f, keep = _format(fmt)
    out = _Layout()
    err = ctypes.create_string_buffer(256)
    _raise(library().benes_hip_layout_of(n, ctypes.byref(f), ctypes.byref(out), err, 256), err)
    d = {name: getattr(out, name) for name, _ in _Layout._fields_
         if name not in ("input_clusters", "output_clusters")}
    d["input_clusters"] = list(out.input_clusters[:out.n_input_clusters])
    d["output_clusters"] = list(out.output_clusters[:out.n_output_clusters])
    return d


def node_address(n: int, fmt: FormatOptions | None, side: int, depth: int, path: int) -> dict:
    f, keep = _format(fmt)
    out = _Address()
    err = ctypes.create_string_buffer(256)
    _raise(library().benes_hip_node_address(n, ctypes.byref(f), side, depth, path,
                                            ctypes.byref(out), err, 256), err)
    return {name: getattr(out, name) for name, _ in _Address._fields_}


def generate(kind: int, n: int, seed: int, rate_ppm: int = 0, width: int = 4) -> np.ndarray:
    """Seeded synthetic permutation (splitmix64 + Lemire bounded draws)."""
    out = np.empty(n, np.uint32 if width == 4 else np.uint64)
    rc = library().benes_hip_generate(kind, n, seed, rate_ppm,
                                      out.ctypes.data_as(ctypes.c_void_p), width)
    if rc != OK:
        raise ValueError("invalid generator arguments")
    return out


def generate_device(kind: int, n: int, seed: int, d_out_ptr: int, width: int = 8,
                    stream: int | None = None) -> None:
    """IDENTITY, REVERSE or KEYED permutation written into device memory."""
    rc = library().benes_hip_generate_device(kind, n, seed, ctypes.c_void_p(d_out_ptr), width,
                                             ctypes.c_void_p(stream or 0))
    if rc != OK:
        raise ValueError("invalid device generator arguments")


def autotune_report() -> str:
    """The subtree-shape tuner's classes, trial device times and choices."""
    lib = library()
    need = lib.benes_hip_autotune_report(None, 0)
    buf = ctypes.create_string_buffer(need + 1)
    lib.benes_hip_autotune_report(buf, need + 1)
    return buf.value.decode()


SUBTREE_SHAPES = ((14, 1024), (13, 1024), (13, 512), (12, 512))
# Calls of one size class that the tuner spends on trials (two per shape);
# BENES_HIP_AUTOTUNE_TRIAL_CALLS in benes_hip.h.
AUTOTUNE_TRIAL_CALLS = 8


def set_subtree_shape(log2_entries: int = 0, threads: int = 0) -> None:
    """Pins a subtree shape (one of SUBTREE_SHAPES); (0, 0) returns to tuning."""
    if library().benes_hip_set_subtree_shape(log2_entries, threads) != OK:
        raise ValueError(f"no subtree shape {log2_entries}x{threads}")


def set_hook_binning(mode: int = -1) -> None:
    """First-hook binning: -1 by size (default), 0 never, 1 wherever a node's
    words exceed 4 MB (the blob is the same in every mode)."""
    if library().benes_hip_set_hook_binning(mode) != OK:
        raise ValueError(f"no hook binning mode {mode}")


def set_wide_words(mode: int = 0) -> None:
    """64-bit union-find words and level-0 inverse: 0 by size (default), 1 at
    every level of every call (the blob is the same in both modes)."""
    if library().benes_hip_set_wide_words(mode) != OK:
        raise ValueError(f"no wide-words mode {mode}")


def set_plain_union(mode: int = -1) -> None:
    """Union pass of the depths whose nodes keep to one XCD: -1 by size
    (default: plain-store hooks, checked and rejoined, where a node's words
    fit in an XCD's L2), 0 compare-and-swap hooks everywhere, 1 plain-store
    hooks at every such depth (the blob is the same in every mode)."""
    if library().benes_hip_set_plain_union(mode) != OK:
        raise ValueError(f"no plain-union mode {mode}")


UNION_FIND, RULERS = 0, 1


def set_colouring(mode: int = UNION_FIND) -> None:
    """Colouring of the HBM levels: UNION_FIND (default) or RULERS, an
    independent second colouring (orbit minima); the blob is the same."""
    if library().benes_hip_set_colouring(mode) != OK:
        raise ValueError(f"no colouring mode {mode}")


def release_workspace() -> None:
    """Frees the library's device buffers for the current device."""
    library().benes_hip_release_workspace()
