# SPDX-License-Identifier: Apache-2.0
"""Benchmark of benes::compress on the gfx950 backend.

Workload (one "step"): compress one uniform Fisher-Yates permutation of
2^26 uint32 entries with middle_group_log2 = 5, input clusters {11, 10},
output clusters {10, 11} and the middle ranks embedded in the input section.
The permutation is resident in HBM before timing starts; each step runs the
whole device path (validation, every routing level, ranks, header) into a
device blob, through the C ABI's device entry point.  The blob is checked
byte for byte against the digest of the reference library's output.

Prints one JSON line.  `roofline` is for the kernel with the largest share of
device time, timed live with HIP events inside the library; `cpu_baseline` is
the reference's own portable_cpu encoder (oracle/_ref, compiled from the
reference sources) on the host cores, on the same input.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "benes-perm_amd"), os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402
import torch  # noqa: E402  (first: libbenes_hip.so then shares torch's HIP runtime)

import benes_hip as B  # noqa: E402

N = 1 << 26
SEED = 0x5EED0000 + N
FMT = B.FormatOptions(5, (11, 10), (10, 11), B.EMBEDDED_INPUT)
HBM_PEAK_GBS = 8000.0


def fixture_digest():
    with open(os.path.join(ROOT, "tests", "golden", "fixtures.json")) as f:
        for rec in json.load(f):
            if rec["name"] == "cfg2-fy-2^26":
                return rec["sha256"]
    return None


def global_levels(n: int, sub: int = 1 << 14) -> int:
    d = 0
    while -(-n // (1 << d)) > sub:
        d += 1
    return d


def algorithmic_bytes(kernel: str, n: int, blob: int) -> float:
    """Bytes one launch of `kernel` must move at minimum: the per-entry
    figures times the entries one launch covers (one launch of a level
    kernel covers one depth of the whole permutation)."""
    per_launch = {
        "k_union": 6.0 * n,             # inverse read (4 B/entry), slot parent words written (4 B/slot)
        "k_route_inputs": 10.0 * n + n / 8,   # P read, children written, slot words read, row bits
        "k_route_outputs": 8.0 * n + n / 4,   # I read, children written, row bits
        "k_level_lds": 16.0 * n + n / 4,      # P and I read, both children written, row bits
        "k_scatter_inverse": 8.0 * n,   # source read, inverse written
        "k_check_inverse": 4.0 * n,
        "k_fill_roots": 4.0 * n,
        "k_pack": n / 8.0,              # one side's row bits of one depth, read and written
    }
    if kernel == "k_subtree":
        lay = B.layout_of(n, FMT)
        switch_bits_above = global_levels(n) * (n // 2) * 2 / 8.0
        return 4.0 * n + (blob - lay["input_offset"] - switch_bits_above)
    return per_launch.get(kernel, 0.0)


def cpu_baseline(perm: np.ndarray, want: str | None):
    import support as S

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    if S.ref_available():
        t0 = time.perf_counter()
        blob = S.ref_compress(perm, 5, B.EMBEDDED_INPUT, (11, 10), (10, 11), threads=cores)
        dt = time.perf_counter() - t0
        ok = want is None or hashlib.sha256(blob).hexdigest() == want
        return {"value": perm.size / dt / 1e6, "unit": "Mentries/s", "cores": cores,
                "kind": "reference",
                "sample": f"full workload (2^26 entries) once, portable_cpu threads={cores}",
                "seconds": dt, "byte_exact": ok}
    sample = 1 << 22
    p = B.generate(B.FISHER_YATES, sample, SEED)
    t0 = time.perf_counter()
    S.oracle_compress(p, 5, B.EMBEDDED_INPUT, (10, 9), (9, 10))
    dt = time.perf_counter() - t0
    return {"value": sample / dt / 1e6, "unit": "Mentries/s", "cores": 1, "kind": "port",
            "sample": "2^22 entries, single-threaded oracle", "seconds": dt}


def calibrate_copy(nbytes: int = 256 << 20, reps: int = 5):
    """A copy of known size, to scale the FETCH_SIZE / WRITE_SIZE counters:
    each launch reads and writes exactly `nbytes`."""
    src = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "device copy", "bytes_read_per_launch": nbytes,
                      "bytes_written_per_launch": nbytes, "launches": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this backend compresses one permutation (or one "
              "batch) per device and has no multi-GPU path; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    torch.cuda.init()
    if args.calibrate_copy:
        calibrate_copy()
        return
    assert B.available(), "gfx950 backend unavailable"
    t0 = time.perf_counter()
    perm = B.generate(B.FISHER_YATES, N, SEED)
    gen_s = time.perf_counter() - t0
    size = B.encoded_size(N, FMT)
    d_perm = torch.from_numpy(perm.view(np.int32)).cuda()
    d_out = torch.empty(size, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        B.compress_device(d_perm.data_ptr(), N, 4, d_out.data_ptr(), size, FMT, stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    device_ms = B.last_device_seconds() * 1e3

    want = fixture_digest()
    blob = d_out.cpu().numpy().tobytes()
    exact = want is not None and hashlib.sha256(blob).hexdigest() == want

    # Per-kernel device time, HIP events on the library's stream.
    B.set_profiling(True)
    prof_steps = 3
    for _ in range(prof_steps):
        step()
    stats = B.kernel_stats()
    B.set_profiling(False)
    total_ms = sum(ms for ms, _ in stats.values())
    dom = max(stats, key=lambda k: stats[k][0])
    dom_ms, dom_launches = stats[dom]
    avg_ms = dom_ms / dom_launches
    alg = algorithmic_bytes(dom, N, size)
    achieved = alg / (avg_ms * 1e-3) / 1e9
    # HBM bytes per launch of the same kernel from the committed counter runs
    # (FETCH_SIZE and WRITE_SIZE, each scaled by a plain copy's reading).
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            rec = json.load(f).get("kernels", {}).get(dom)
        if rec:
            traffic = rec["bytes_per_launch"]

    ms_per_step = elapsed / args.steps * 1e3
    line = {
        "metric": "Mentries/s compressed (byte-exact), plain compress 2^26 uint32",
        "value": N * args.steps / elapsed / 1e6,
        "unit": "Mentries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "plain compress, one uniform Fisher-Yates permutation of 2^26 uint32",
            "n": N, "seed": SEED, "middle_group_log2": 5, "input_clusters": [11, 10],
            "output_clusters": [10, 11], "middle_placement": "embedded_input",
            "blob_bytes": size,
        },
        "byte_exact": exact,
        "device_ms": device_ms,
        "roofline": {
            "bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "algorithmic_bytes_per_launch": alg, "avg_launch_ms": avg_ms,
            "launches_per_step": dom_launches / prof_steps,
        },
        "kernel_ms_per_step": {k: v[0] / prof_steps for k, v in sorted(stats.items())},
        "pipeline_roofline": {
            "algorithmic_bytes": 4 * N + 8 * N * 11 + size,
            "kernel_ms": total_ms / prof_steps,
            "frac_of_8TBs": (4 * N + 8 * N * 11 + size) / (ms_per_step * 1e-3) / 8e12,
        },
        "setup": {"generate_s": gen_s},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(perm, want)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
