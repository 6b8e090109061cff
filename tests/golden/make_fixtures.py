# SPDX-License-Identifier: Apache-2.0
"""Writes tests/golden/fixtures.json: SHA-256 digests of blobs produced by the
reference library itself (oracle/_ref/libbenes_ref.so, compiled from the
reference's src/benes.cpp and src/backends_stub.cpp by oracle/Makefile) on
seeded inputs from benes_hip.generate().

Run in a container where the reference is mounted and built:
    make -C oracle && make -C benes-perm_amd && python tests/golden/make_fixtures.py
(`--structured` rewrites only tests/golden/structured.json, the digests of
the structured inputs of support.structured).

Each record holds the input recipe (generator kind, n, seed, rate, width) and
the format, so the same input can be rebuilt anywhere without the reference.
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import support as S  # noqa: E402

B = S.benes_hip

CFG2_FMT = dict(g=5, placement=1, ic=[11, 10], oc=[10, 11])


def cases():
    out = []
    # Plain compress, 2^20 uint32, default format (the reference's CPU case).
    out.append(dict(name="cfg1-fy-2^20", kind=B.FISHER_YATES, n=1 << 20, seed=0x5EED0000 + (1 << 20),
                    rate=0, width=4, g=5, placement=0, ic=[], oc=[]))
    # 2^26 uint32 with input {11,10}, output {10,11}, embedded_input.
    n = 1 << 26
    out.append(dict(name="cfg2-fy-2^26", kind=B.FISHER_YATES, n=n, seed=0x5EED0000 + n, rate=0,
                    width=4, **CFG2_FMT))
    out.append(dict(name="cfg2-cycle-2^26", kind=B.SINGLE_CYCLE, n=n, seed=0x5EED1000 + n, rate=0,
                    width=4, **CFG2_FMT))
    out.append(dict(name="cfg2-nearid-2^26", kind=B.NEAR_IDENTITY, n=n, seed=0x5EED2000 + n,
                    rate=10000, width=4, **CFG2_FMT))
    # Boundary sizes around powers of two.
    for n in (4095, 4097, 65535, 65537, (1 << 20) - 1, (1 << 20) + 1):
        out.append(dict(name=f"fy-{n}", kind=B.FISHER_YATES, n=n, seed=0xB0DA0000 + n, rate=0,
                        width=4, g=5, placement=0, ic=[], oc=[]))
    # 64-bit source, non-power-of-two, embedded_output with default clusters.
    for n in ((1 << 20) + 12345, (1 << 22) + 12345):
        out.append(dict(name=f"u64-fy-{n}", kind=B.FISHER_YATES, n=n, seed=0x64640000 + n, rate=0,
                        width=8, g=5, placement=2, ic=[], oc=[]))
    # Batched-style items: g = 4, single cycles and near-identity.
    for n, kind, rate in ((1000, B.SINGLE_CYCLE, 0), (3001, B.NEAR_IDENTITY, 5000),
                          (150001, B.FISHER_YATES, 0), (262144 + 77, B.SINGLE_CYCLE, 0)):
        out.append(dict(name=f"g4-{kind}-{n}", kind=kind, n=n, seed=0xBA7C0000 + n, rate=rate,
                        width=4, g=4, placement=0, ic=[], oc=[]))
    # Multi-cluster, tail-packet heavy formats on pruned sizes.
    for n, g, pl, ic, oc in ((100003, 3, 1, [2, 5, 7], [7, 5, 2]),
                             (77777, 5, 2, [1] * 12, [12]),
                             (4097, 4, 1, [3, 6], [5, 4])):
        out.append(dict(name=f"fmt-{n}-{g}-{pl}", kind=B.FISHER_YATES, n=n, seed=0xF0F00000 + n,
                        rate=0, width=4, g=g, placement=pl, ic=ic, oc=oc))
    return out


def structured_cases():
    """Structured inputs (support.structured) at the 2^26 scale, whose level
    graphs are long chains for the union-find."""
    small = [dict(name=f"{kind}-{n}", kind=kind, n=n, width=4, g=5, placement=pl, ic=[], oc=[])
             for kind, n, pl in (("reverse", 65537, 0), ("rotate1", 65536, 1), ("rotate3", 65536, 2),
                                 ("rotate_half", 131072, 0), ("shuffle", 65537, 1),
                                 ("reverse_halves", 131073, 2))]
    return small + [
        dict(name="rotate1-2^26", kind="rotate1", n=1 << 26, width=4, **CFG2_FMT),
        dict(name="reverse-2^26+1", kind="reverse", n=(1 << 26) + 1, width=4, g=5, placement=1,
             ic=[], oc=[]),
        dict(name="shuffle-2^26+1", kind="shuffle", n=(1 << 26) + 1, width=4, g=5, placement=1,
             ic=[], oc=[]),
    ]


def write_structured():
    records = []
    for c in structured_cases():
        perm = S.structured(c["kind"], c["n"])
        blob = S.ref_compress(perm, c["g"], c["placement"], c["ic"], c["oc"], threads=8)
        rec = dict(c)
        rec["size"] = len(blob)
        rec["sha256"] = hashlib.sha256(blob).hexdigest()
        records.append(rec)
        print(rec["name"], rec["size"], rec["sha256"][:16], flush=True)
    with open(os.path.join(HERE, "structured.json"), "w") as f:
        json.dump(records, f, indent=1)


def main():
    if not S.ref_available():
        raise SystemExit("oracle/_ref/libbenes_ref.so is not built (needs the reference sources)")
    records = []
    for c in cases():
        perm = B.generate(c["kind"], c["n"], c["seed"], c["rate"], c["width"])
        blob = S.ref_compress(perm, c["g"], c["placement"], c["ic"], c["oc"], threads=8)
        rec = dict(c)
        rec["size"] = len(blob)
        rec["sha256"] = hashlib.sha256(blob).hexdigest()
        rec["first8"] = [int(x) for x in perm[:8]]
        records.append(rec)
        print(rec["name"], rec["size"], rec["sha256"][:16], flush=True)
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(records, f, indent=1)
    write_structured()


if __name__ == "__main__":
    if sys.argv[1:] == ["--structured"]:
        if not S.ref_available():
            raise SystemExit("oracle/_ref/libbenes_ref.so is not built (needs the reference sources)")
        write_structured()
    else:
        main()
