# SPDX-License-Identifier: Apache-2.0
"""The allocation-free entry point benes_hip_compress_device_noalloc, called
through the C ABI, against benes_hip_compress_device on the same device
inputs: two calls in a row, the second with a smaller input (so the shared
scratch holds the first call's contents), both widths, the largest size it
takes, and its refusals.  It has no counterpart in the reference; its bytes
must be those of the existing device entry point."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import support as S

B = S.benes_hip
pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def device():
    assert B.available(), "no gfx950 device or libbenes_hip.so failed to load its kernels"


def _device_perm(kind, n, seed, width):
    perm = B.generate(kind, n, seed, 5000, width)
    view = perm.view(np.int32 if width == 4 else np.int64)
    return torch.from_numpy(view).cuda()


def _noalloc(d_perm, n, width, fmt, d_out, capacity, stream):
    """Direct ctypes call of the C ABI entry point; returns (status, size, message)."""
    lib = B.library()
    f, keep = B._format(fmt)
    size = ctypes.c_uint64()
    err = ctypes.create_string_buffer(256)
    rc = lib.benes_hip_compress_device_noalloc(
        ctypes.c_void_p(d_perm), n, width, ctypes.byref(f), ctypes.c_void_p(d_out), capacity,
        ctypes.byref(size), ctypes.c_void_p(stream), err, 256)
    del keep
    return rc, size.value, err.value.decode()


def _both(kind, n, seed, width, fmt):
    stream = torch.cuda.current_stream().cuda_stream
    d_perm = _device_perm(kind, n, seed, width)
    size = B.encoded_size(n, fmt)
    cap = (size + 7) // 8 * 8
    want = torch.zeros(cap, dtype=torch.uint8, device="cuda")
    got = torch.full((cap,), 0xA5, dtype=torch.uint8, device="cuda")
    assert B.compress_device(d_perm.data_ptr(), n, width, want.data_ptr(), cap, fmt,
                             stream) == size
    print(f"noalloc case kind={kind} n={n} width={width}", flush=True)
    rc, got_size, msg = _noalloc(d_perm.data_ptr(), n, width, fmt, got.data_ptr(), cap, stream)
    assert rc == B.OK, msg
    assert got_size == size
    torch.cuda.synchronize()
    assert torch.equal(got[:size], want[:size]), (kind, n, width)
    return got[:size].cpu().numpy().tobytes()


def test_two_calls_second_smaller_match_device_entry_point():
    fmt1 = B.FormatOptions(5, (8, 7), (10, 5), B.EMBEDDED_INPUT)
    _both(B.FISHER_YATES, B.NOALLOC_MAX_ENTRIES, 11, 4, fmt1)
    fmt2 = B.FormatOptions(4, (), (), B.EMBEDDED_OUTPUT)
    blob = _both(B.SINGLE_CYCLE, 100003, 12, 8, fmt2)
    perm = B.generate(B.SINGLE_CYCLE, 100003, 12, 5000, 8)
    assert blob == S.oracle_compress(perm, 4, B.EMBEDDED_OUTPUT)


def test_sizes_shapes_and_widths():
    # Small to the largest size, every subtree shape (the deepest plan is the
    # smallest shape's), 64-bit words forced (the most scratch).
    cases = [(B.FISHER_YATES, 1, 4), (B.NEAR_IDENTITY, 33, 8), (B.FISHER_YATES, 65537, 4),
             (B.FISHER_YATES, (1 << 20) - 3, 8)]
    for kind, n, width in cases:
        _both(kind, n, n, width, B.FormatOptions())
    for shape in B.SUBTREE_SHAPES:
        B.set_subtree_shape(*shape)
        try:
            _both(B.FISHER_YATES, B.NOALLOC_MAX_ENTRIES, 77, 4, B.FormatOptions(5, (), (), 0))
        finally:
            B.set_subtree_shape(0, 0)
    B.set_wide_words(1)
    B.set_subtree_shape(12, 512)
    try:
        _both(B.FISHER_YATES, B.NOALLOC_MAX_ENTRIES, 78, 8, B.FormatOptions(5, (), (), 2))
    finally:
        B.set_wide_words(0)
        B.set_subtree_shape(0, 0)


def test_refusals_and_recovery():
    stream = torch.cuda.current_stream().cuda_stream
    n = B.NOALLOC_MAX_ENTRIES + 1
    d_perm = _device_perm(B.IDENTITY, n, 0, 4)
    size = B.encoded_size(n, B.FormatOptions())
    d_out = torch.zeros((size + 7) // 8 * 8, dtype=torch.uint8, device="cuda")
    rc, got_size, msg = _noalloc(d_perm.data_ptr(), n, 4, B.FormatOptions(), d_out.data_ptr(),
                                 d_out.numel(), stream)
    assert rc == B.LENGTH and "at most 2^20 entries" in msg and got_size == size
    # Output too small for whole words, and misaligned.
    n = 1000
    d_perm = _device_perm(B.FISHER_YATES, n, 3, 4)
    size = B.encoded_size(n, B.FormatOptions())
    d_out = torch.zeros((size + 7) // 8 * 8 + 8, dtype=torch.uint8, device="cuda")
    rc, got_size, _ = _noalloc(d_perm.data_ptr(), n, 4, B.FormatOptions(), d_out.data_ptr(),
                               size - 1, stream)
    assert rc == B.BUFFER_TOO_SMALL and got_size == size
    rc, _, _ = _noalloc(d_perm.data_ptr(), n, 4, B.FormatOptions(), d_out.data_ptr() + 4,
                        d_out.numel() - 4, stream)
    assert rc == B.BUFFER_TOO_SMALL
    # Not a permutation: the reference's message; the next call is still right.
    bad = torch.zeros(n, dtype=torch.int32, device="cuda")
    rc, _, msg = _noalloc(bad.data_ptr(), n, 4, B.FormatOptions(), d_out.data_ptr(),
                          d_out.numel(), stream)
    assert rc == B.INVALID_ARGUMENT and "every value from zero" in msg
    rc, _, msg = _noalloc(d_perm.data_ptr(), n, 4, B.FormatOptions(), d_out.data_ptr(),
                          d_out.numel(), stream)
    assert rc == B.OK, msg
    want = S.oracle_compress(B.generate(B.FISHER_YATES, n, 3, 5000, 4))
    assert d_out[:size].cpu().numpy().tobytes() == want
