# SPDX-License-Identifier: Apache-2.0
"""The 64-bit path at its full size: N = 2^32 + 12345 uint64 entries,
middle_group_log2 5, default clusters, ranks embedded in the output section.

At this size the root has more than 2^31 input slots and the inverse needs
64-bit words, so the wide union-find and inverse variants run, along with the
pruned ceil/floor recursion and the short final groups.  An oracle encode at
this size is out of reach, so the blobs are checked through properties that
do not depend on size:
  * reverse: a reversal routes into two reversals; every input switch and
    every rank is 0, and the output switches are 1 exactly at nodes of even
    size (the closed form the oracle tests pin against the reference's
    reverse-33 golden blob and the oracle's own encoder), so the input section
    is all zero and the output section holds exactly the even nodes' pairs as
    ones;
  * identity: every switch bit is 0 and every leaf rank is c! - 1, so the
    input section is all zero, the output section's ones are exactly those of
    the ranks, and sampled entries decode to themselves;
  * a keyed pseudo-random bijection: sampled entries decode, forward and
    backward, to the values that were compressed (the oracle's decoder), and
    the device lookup decodes all 4.3e9 entries back to the input;
  * a uniformly random permutation: its blob equals, byte for byte, the blob
    of a second colouring that shares no code with the union-find (the
    rulers), which checks every switch bit as canonical, and sampled entries
    decode.
The structured and keyed inputs are generated on the device (the
permutation, 34 GB, never crosses to the host), the uniform one on the
host's threads; the blob (16.5 GB) crosses to the host once per case for the
oracle's decoder.
"""
from __future__ import annotations

import functools
import math
import sys
import time

import numpy as np
import pytest

import support as S

B = S.benes_hip
pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N = (1 << 32) + 12345
FMT = S.fmt(5, 2)          # default clusters {14,14}/{14,14}, embedded_output
SAMPLES = 3000


@pytest.fixture(scope="module")
def big():
    assert B.available(), "no gfx950 device or libbenes_hip.so failed to load its kernels"
    free, _ = torch.cuda.mem_get_info()
    if free < (200 << 30):
        pytest.skip(f"needs about 200 GB of free device memory, {free >> 30} GB free")
    size = B.encoded_size(N, FMT)
    d_perm = torch.empty(N, dtype=torch.int64, device="cuda")
    d_out = torch.empty((size + 7) // 8 * 8, dtype=torch.uint8, device="cuda")
    lay = B.layout_of(N, FMT)
    yield d_perm, d_out, size, lay
    del d_perm, d_out
    torch.cuda.empty_cache()
    B.release_workspace()


LEAF = 32


@functools.lru_cache(maxsize=None)
def even_pairs(c: int) -> int:
    """Output switches set by a reversal of c entries (pairs of even nodes)."""
    if c <= LEAF:
        return 0
    return (c // 2 if c % 2 == 0 else 0) + even_pairs((c + 1) // 2) + even_pairs(c // 2)


@functools.lru_cache(maxsize=None)
def identity_rank_ones(c: int) -> int:
    """Ones in the ranks of an identity of c entries (c! - 1 at every leaf)."""
    if c <= LEAF:
        return bin(math.factorial(c) - 1).count("1")
    return identity_rank_ones((c + 1) // 2) + identity_rank_ones(c // 2)


def ones(t) -> int:
    """Set bits of a uint8 device tensor, counted in slices."""
    total = 0
    step = 256 << 20
    for a in range(0, t.numel(), step):
        x = t[a:a + step].to(torch.int32)
        for k in range(8):
            total += int(((x >> k) & 1).sum())
    return total


def progress(msg: str) -> None:
    print(f"[large] {msg}", file=sys.stderr, flush=True)


def run(big, kind, seed=0):
    d_perm, d_out, size, _ = big
    progress(f"kind {kind}: generating {N} entries")
    B.generate_device(kind, N, seed, d_perm.data_ptr(), 8)
    torch.cuda.synchronize()
    progress(f"kind {kind}: compressing")
    t0 = time.perf_counter()
    got = B.compress_device(d_perm.data_ptr(), N, 8, d_out.data_ptr(), d_out.numel(), FMT)
    torch.cuda.synchronize()
    progress(f"kind {kind}: compressed in {time.perf_counter() - t0:.1f} s")
    assert got == size
    return d_out[:size]


def sample_indices(seed):
    rng = np.random.default_rng(seed)
    edges = np.array([0, 1, 2, N // 2, N - 3, N - 2, N - 1], np.uint64)
    return np.concatenate([edges, rng.integers(0, N, SAMPLES, dtype=np.uint64)])


def test_generators_agree_with_host():
    n = (1 << 20) + 7
    d = torch.empty(n, dtype=torch.int64, device="cuda")
    for kind in (B.IDENTITY, B.REVERSE, B.KEYED):
        B.generate_device(kind, n, 99, d.data_ptr(), 8)
        assert np.array_equal(d.cpu().numpy().view(np.uint64), B.generate(kind, n, 99, 0, 8))


def test_reverse_closed_form(big):
    blob = run(big, B.REVERSE)
    lay = big[3]
    assert int(torch.count_nonzero(blob[lay["input_offset"]:lay["output_offset"]])) == 0
    assert ones(blob[lay["output_offset"]:]) == even_pairs(N)
    host = blob[:4096].cpu().numpy()
    assert int.from_bytes(host[8:16].tobytes(), "little") == N   # entry count in the header


def test_identity(big):
    blob = run(big, B.IDENTITY)
    lay = big[3]
    assert int(torch.count_nonzero(blob[lay["input_offset"]:lay["output_offset"]])) == 0
    assert ones(blob[lay["output_offset"]:]) == identity_rank_ones(N)
    host = blob.cpu().numpy()
    progress("identity: blob copied to the host")
    idx = sample_indices(1)
    assert np.array_equal(S.oracle_lookup(host, idx), idx)
    assert np.array_equal(S.oracle_lookup(host, idx, inverse=True), idx)
    progress("identity: sampled lookups done")


def test_keyed_round_trip(big):
    blob = run(big, B.KEYED, seed=0x5EED)
    d_perm = big[0]
    host = blob.cpu().numpy()
    progress("keyed: blob copied to the host")
    idx = sample_indices(2)
    want = d_perm[torch.from_numpy(idx.astype(np.int64)).cuda()].cpu().numpy().view(np.uint64)
    assert np.array_equal(S.oracle_lookup(host, idx), want)
    assert np.array_equal(S.oracle_lookup(host, want, inverse=True), idx)
    progress("keyed: sampled lookups done")
    # And every entry: the device lookup (checked against the oracle's decoder
    # in test_gpu_lookup.py) decodes the whole blob back to the input.
    del host
    B.release_workspace()
    d_dec = torch.empty(N, dtype=torch.int64, device="cuda")
    B.lookup_device(blob.data_ptr(), blob.numel(), None, N, d_dec.data_ptr(), 8)
    torch.cuda.synchronize()
    progress("keyed: whole blob decoded on the device")
    assert torch.equal(d_dec, d_perm)
    del d_dec


def equal_in_slices(a, b, step=1 << 30) -> bool:
    return all(torch.equal(a[i:i + step], b[i:i + step]) for i in range(0, a.numel(), step))


def test_uniform_two_colourings(big):
    """A uniformly random permutation of the full size (drawn on the host's
    threads by Rao-Sandelius bucketing and per-bucket Fisher-Yates: every
    arrangement equally likely, as a seeded Fisher-Yates draw would give),
    compressed with the union-find colouring and again with the ruler
    colouring, which shares no code with it and is pinned against the oracle
    and the reference digests by tests/test_gpu_rulers.py.  The two blobs
    must be byte-identical: every switch bit of every level, coloured by the
    64-bit union-find at the root and the 32-bit one below, is the canonical
    one.  Sampled entries then decode through the oracle's decoder."""
    d_perm, d_out, size, _ = big
    progress("uniform: drawing the permutation on the host")
    t0 = time.perf_counter()
    host = B.generate(B.UNIFORM_PARALLEL, N, 0x5EED0000 + 32, 0, 8)
    progress(f"uniform: drawn in {time.perf_counter() - t0:.1f} s")
    d_perm.copy_(torch.from_numpy(host.view(np.int64)))
    torch.cuda.synchronize()
    progress("uniform: compressing with the union-find")
    assert B.compress_device(d_perm.data_ptr(), N, 8, d_out.data_ptr(), d_out.numel(), FMT) == size
    torch.cuda.synchronize()
    d_ref = torch.empty_like(d_out)
    B.set_colouring(B.RULERS)
    try:
        progress("uniform: compressing with the rulers")
        t0 = time.perf_counter()
        assert B.compress_device(d_perm.data_ptr(), N, 8, d_ref.data_ptr(), d_ref.numel(),
                                 FMT) == size
        torch.cuda.synchronize()
        progress(f"uniform: ruler compress {time.perf_counter() - t0:.1f} s")
    finally:
        B.set_colouring(B.UNION_FIND)
    same = equal_in_slices(d_out[:size], d_ref[:size])
    del d_ref
    assert same, "union-find and ruler colourings differ at 2^32 + 12345"
    progress("uniform: blobs identical")
    blob = d_out[:size].cpu().numpy()
    idx = sample_indices(3)
    assert np.array_equal(S.oracle_lookup(blob, idx), host[idx])
    assert np.array_equal(S.oracle_lookup(blob, host[idx], inverse=True), idx)
    progress("uniform: sampled lookups done")
