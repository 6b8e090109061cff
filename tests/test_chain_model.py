# SPDX-License-Identifier: Apache-2.0
"""CPU model of one HBM level's union-find passes (k_hook_first, then the
Rem walks of k_union, run one edge at a time), on a whole node.

It measures what bounds k_union's critical path: the longest chain of
dependent parent loads a single edge's walk follows.  On random permutations
that is a few dozen steps; on structured ones the first-hook pass hangs every
slot under its neighbour and the edge closing the cycle walks a quarter to a
half of the node, which is why k_union caps its walks (kWalkCap) and defers
the rest to the flatten pass.  The model also checks that the forest it
builds colours every slot as the canonical rule does (the odd tail, then each
component's smallest slot, takes colour 0), the rule itself being pinned
against the oracle's input switch bits of one-level permutations."""
from __future__ import annotations

import numpy as np
import pytest

import support as S

ROOT = -1


def in_slot(x, cnt):
    if cnt & 1:
        return np.where(x == cnt - 1, 0, (x >> 1) + 1)
    return x >> 1


def two_passes(perm: np.ndarray):
    """Forest (parent, parity) after both passes, and the longest Rem walk."""
    n = perm.size
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    h = n >> 1
    a, b = inv[0:2 * h:2], inv[1:2 * h:2]
    sa, sb = in_slot(a, n), in_slot(b, n)
    lab = 1 ^ ((a ^ b) & 1)
    lc = h + (n & 1)
    par = np.full(lc, ROOT, np.int64)
    pbit = np.zeros(lc, np.int64)
    for e, s, o in ((a, sa, sb), (b, sb, sa)):     # k_hook_first
        m = ((e & 1) == 0) & (o < s)
        par[s[m]] = o[m]
        pbit[s[m]] = lab[m]
    lo, hi = np.minimum(sa, sb), np.maximum(sa, sb)
    hi_odd = np.where(sa < sb, b, a) & 1
    longest = 0
    for q in np.nonzero((lo != hi) & (hi_odd == 1))[0]:   # k_union
        l, hh, lb = int(lo[q]), int(hi[q]), int(lab[q])
        if par[hh] == ROOT:
            par[hh], pbit[hh] = l, lb
            continue
        x, px, y, py, steps = int(par[hh]), int(pbit[hh]), l, 0, 0
        while x != y:
            if x < y:
                x, y, px, py = y, x, py, px
            steps += 1
            if par[x] == ROOT:
                par[x], pbit[x] = y, (px ^ py ^ lb) & 1
                break
            px ^= int(pbit[x])
            x = int(par[x])
        longest = max(longest, steps)
    return par, pbit, longest


def colours(par, pbit):
    out = np.zeros(par.size, np.int64)
    for s in range(par.size):
        c, x = 0, s
        while par[x] != ROOT:
            c ^= int(pbit[x])
            x = int(par[x])
        out[s] = c
    return out


def canonical(perm: np.ndarray) -> np.ndarray:
    """Switch value of every input slot by the reference's rule: colour the
    odd tail 0, then each uncoloured slot in ascending order 0, propagating
    'different colour' along output pairs (src/benes.cpp:278-297)."""
    n = perm.size
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    colour = np.full(n, -1, np.int64)
    partner_out = np.empty(n, np.int64)          # entry -> the other entry of its output pair
    for q in range(n >> 1):
        partner_out[inv[2 * q]], partner_out[inv[2 * q + 1]] = inv[2 * q + 1], inv[2 * q]
    order = ([n - 1] if n & 1 else []) + list(range(n - (n & 1)))
    for start in order:
        if colour[start] >= 0:
            continue
        stack = [(start, 0)]
        while stack:
            x, c = stack.pop()
            if colour[x] >= 0:
                continue
            colour[x] = c
            if not (n & 1 and x == n - 1):
                stack.append((x ^ 1, 1 - c))          # the input pair's other entry
            if not (n & 1 and perm[x] == n - 1):
                stack.append((int(partner_out[x]), 1 - c))
    slots = np.arange(n >> 1) + (n & 1)
    out = np.zeros((n >> 1) + (n & 1), np.int64)
    out[slots] = colour[2 * np.arange(n >> 1)]
    return out


@pytest.mark.parametrize("n", [64, 65, 1000, 1001])
def test_forest_colours_canonically(n):
    for seed in range(3):
        perm = np.random.default_rng(seed).permutation(n)
        par, pbit, _ = two_passes(perm)
        assert np.array_equal(colours(par, pbit), canonical(perm))
    for kind in ("reverse", "rotate1", "shuffle", "reverse_halves"):
        perm = S.structured(kind, n).astype(np.int64)
        par, pbit, _ = two_passes(perm)
        assert np.array_equal(colours(par, pbit), canonical(perm)), kind


@pytest.mark.parametrize("n", [33, 47, 64])
def test_canonical_rule_matches_the_oracle(n):
    # One level of switches (g = 5, 32 < n <= 64): the input section holds
    # colour(2p) for every pair p, LSB first.
    for seed in range(5):
        perm = np.random.default_rng(100 + seed).permutation(n)
        blob = np.frombuffer(S.oracle_compress(perm.astype(np.uint64), 5, 0), np.uint8)
        lay = S.oracle_layout(n, 5, 0)
        bits = np.unpackbits(blob[lay.input_offset:lay.output_offset], bitorder="little")
        want = canonical(perm)[np.arange(n >> 1) + (n & 1)]
        assert np.array_equal(bits[: n >> 1], want)


def test_random_walks_are_short():
    for seed in range(4):
        _, _, longest = two_passes(np.random.default_rng(seed).permutation((1 << 16) + seed))
        assert longest < 64


@pytest.mark.parametrize("kind,n", [("reverse", (1 << 16) + 1), ("rotate1", 1 << 16),
                                    ("shuffle", (1 << 16) + 1)])
def test_structured_walks_span_the_node(kind, n):
    _, _, longest = two_passes(S.structured(kind, n).astype(np.int64))
    assert longest >= n // 8 > 1024   # far beyond k_union's walk cap
