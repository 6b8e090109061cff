# SPDX-License-Identifier: Apache-2.0
"""CPU model of the ruler colouring (k_ruler_walk, k_ruler_jump,
k_ruler_route), the second colouring of the HBM levels that shares no code
with the union-find: colour(x) is the parity of the smallest entry of x's
orbit under x -> I[P[x] ^ 1] ^ 1, or the end its chain reaches in an odd
node.  The model runs the kernels' steps one thread at a time (same ruler
hash, same end markers, same number of jumping rounds) and checks the input
switch values against the reference's canonical rule (the odd tail, then each
component's smallest input, takes colour 0; src/benes.cpp:278-297, restated
by tests/test_chain_model.py and pinned there against the oracle), on random,
structured and rulerless inputs."""
from __future__ import annotations

import numpy as np
import pytest

import support as S
from test_chain_model import canonical

LOG2 = 5
END_A, END_B, UNWALKED, UNOWNED = 0xFFFFFFFF, 0xFFFFFFFE, 0xFFFFFFFD, 0xFFFFFFFF
M64 = (1 << 64) - 1


def is_ruler(a: int) -> bool:
    z = ((a >> LOG2) * 0x9E3779B97F4A7C15) & M64
    z ^= z >> 29
    z = (z * 0xBF58476D1CE4E5B9) & M64
    z ^= z >> 32
    return (a & 31) == (z & 0xFFFFFFFF & 31)


def ruler_switches(perm: np.ndarray, off: int = 0):
    """Input switch values s_p of one node whose entries sit at [off, off + c)."""
    c = perm.size
    P = [int(v) for v in perm]
    inv = [0] * c
    for x, y in enumerate(P):
        inv[y] = x
    odd = c & 1

    def nxt(x):
        y = P[x]
        if odd and y == c - 1:
            return "A"
        z = inv[y ^ 1]
        if odd and z == c - 1:
            return "B"
        return z ^ 1

    buckets = (off + c >> LOG2) + 2
    owner = {}
    rnext = [UNWALKED] * buckets
    rmin = [0] * buckets
    walked = 0
    for x in range(c):                        # k_ruler_walk
        if not is_ruler(off + x):
            continue
        rid = (off + x) >> LOG2
        cur, m = x, x
        while True:
            v = nxt(cur)
            walked += 1
            if v in ("A", "B"):
                n = END_A if v == "A" else END_B
                break
            if is_ruler(off + v):
                n = (off + v) >> LOG2
                break
            owner[v] = rid
            m = min(m, v)
            cur = v
        owner[x] = rid
        rnext[rid], rmin[rid] = n, m
    rounds = 0
    while (1 << rounds) < -(-c // 32) + 1:
        rounds += 1
    for _ in range(rounds):                    # k_ruler_jump, synchronous
        nn, mm = rnext[:], rmin[:]
        for i in range(buckets):
            j = rnext[i]
            if j >= UNWALKED:
                continue
            nn[i], mm[i] = rnext[j], min(rmin[i], rmin[j])
        rnext, rmin = nn, mm

    def rcol(i):
        return 0 if rnext[i] == END_A else 1 if rnext[i] == END_B else rmin[i] & 1

    s = []
    orphans = 0
    for p in range(c >> 1):                    # k_ruler_route
        x = 2 * p
        if is_ruler(off + x):
            s.append(rcol((off + x) >> LOG2))
            continue
        if x in owner:
            s.append(rcol(owner[x]))
            continue
        orphans += 1
        cur, m = x, x
        while True:
            v = nxt(cur)
            if v in ("A", "B"):
                s.append(0 if v == "A" else 1)
                break
            if v == x:
                s.append(m & 1)
                break
            if is_ruler(off + v):
                s.append(rcol((off + v) >> LOG2))
                break
            if v in owner:
                s.append(rcol(owner[v]))
                break
            m = min(m, v)
            cur = v
    return np.array(s, np.int64), walked, orphans


def check(perm, off=0):
    s, _, orphans = ruler_switches(perm, off)
    n = perm.size
    want = canonical(perm)[np.arange(n >> 1) + (n & 1)]
    assert np.array_equal(s, want)
    return orphans


@pytest.mark.parametrize("n", [2, 3, 33, 64, 65, 1000, 1001, 4097])
def test_random_matches_canonical(n):
    for seed in range(3):
        check(np.random.default_rng(seed).permutation(n), off=seed * 7919)


@pytest.mark.parametrize("kind", ["reverse", "rotate1", "rotate3", "rotate_half", "shuffle",
                                  "reverse_halves"])
@pytest.mark.parametrize("n", [1024, 1025, 3001])
def test_structured_matches_canonical(kind, n):
    check(S.structured(kind, n).astype(np.int64), off=n)


def test_rulerless_orbits():
    # The identity: every orbit is one entry, almost all without a ruler.
    assert check(np.arange(2000), off=13) > 900
    # Many short cycles: products of transpositions and 3-cycles.
    rng = np.random.default_rng(3)
    perm = np.arange(3000)
    for i in range(0, 2990, 7):
        perm[[i, i + 3]] = perm[[i + 3, i]]
        perm[[i + 1, i + 2, i + 5]] = perm[[i + 2, i + 5, i + 1]]
    check(perm, off=int(rng.integers(1 << 20)))


def test_every_orbit_sampled_once():
    # Ruler walks cover the ruler-bearing orbits once: total steps about n.
    n = 1 << 14
    _, walked, orphans = ruler_switches(np.random.default_rng(9).permutation(n))
    assert n // 2 <= walked <= n + 64 and orphans < 64
