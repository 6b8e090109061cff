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


def fixture(name: str):
    with open(os.path.join(ROOT, "tests", "golden", "fixtures.json")) as f:
        for rec in json.load(f):
            if rec["name"] == name:
                return rec
    return None


def fixture_digest(name: str = "cfg2-fy-2^26"):
    rec = fixture(name)
    return rec["sha256"] if rec else None


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
        "k_hook_first": 6.0 * n,        # inverse read (4 B/entry), every slot's parent word written (4 B/slot)
        "k_union": 4.0 * n,             # inverse read
        "k_union_plain": 4.0 * n,       # inverse read (the same pass with plain-store hooks)
        "k_route_inputs": 10.0 * n + n / 8,   # P read, children written, slot words read, row bits
        "k_route_outputs": 8.0 * n + n / 4,   # I read, children written, row bits
        "k_level_lds": 8.0 * n + n / 4,       # P and I read, both children written (16-bit entries), row bits
        "k_scatter_inverse": 8.0 * n,   # source read, inverse written
        "k_check_inverse": 4.0 * n,
        "k_pack": n / 8.0,              # one side's row bits of one depth, read and written
    }
    if kernel == "k_subtree":
        lay = B.layout_of(n, FMT)
        switch_bits_above = global_levels(n) * (n // 2) * 2 / 8.0
        return 4.0 * n + (blob - lay["input_offset"] - switch_bits_above)
    return per_launch.get(kernel, 0.0)


def union_pass(stats: dict, prof_steps: int) -> dict:
    kernels = ["k_union", "k_union_plain", "k_union_verify", "k_union_rejoin"]
    ms = sum(stats.get(k, (0.0, 0))[0] for k in kernels)
    depths = (stats.get("k_union", (0, 0))[1] + stats.get("k_union_plain", (0, 0))[1]) / prof_steps
    per_depth = ms / prof_steps / depths if depths else 0.0
    achieved = 4.0 * N / (per_depth * 1e-3) / 1e9 if per_depth else 0.0
    return {"kernels": kernels,
            "ms_per_step": ms / prof_steps, "depths": depths, "ms_per_depth": per_depth,
            "achieved_GBs": achieved, "frac": achieved / HBM_PEAK_GBS}


def _median_seconds(fn, runs: int = 5):
    """Median wall time of `runs` calls (steady clock around each call), as
    the reference's benchmark times compress (tools/benchmark.cpp:54-68)."""
    times, result = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        result = fn()
        times.append(time.perf_counter() - t0)
    return sorted(times)[len(times) // 2], times, result


def cpu_baseline(runs: int = 5):
    """The reference's portable_cpu encoder (oracle/_ref, compiled from the
    reference sources) on this host's cores, median of `runs` runs each:
    a 2^24-entry sample of the headline workload (the line's value), and the
    2^20 default-format configuration on one thread and on every core."""
    import support as S

    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    if not S.ref_available():
        sample = 1 << 22
        p = B.generate(B.FISHER_YATES, sample, SEED)
        dt, times, _ = _median_seconds(
            lambda: S.oracle_compress(p, 5, B.EMBEDDED_INPUT, (10, 9), (9, 10)), runs)
        return {"value": sample / dt / 1e6, "unit": "Mentries/s", "cores": 1, "kind": "port",
                "sample": f"2^22 entries, single-threaded oracle, median of {runs}",
                "seconds": times}
    # The headline workload's format scaled to 2^24 entries (19 routing
    # levels: input clusters {10, 9}, output {9, 10}, ranks in the input section).
    sample = 1 << 24
    p = B.generate(B.FISHER_YATES, sample, 0x5EED0000 + sample)
    dt, times, _ = _median_seconds(
        lambda: S.ref_compress(p, 5, B.EMBEDDED_INPUT, (10, 9), (9, 10), threads=cores), runs)
    del p
    # Config 1: 2^20 entries, default format, seed 0x5EED0000 + n.
    rec = fixture("cfg1-fy-2^20")
    p1 = B.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"], rec["width"])
    cfg1 = {}
    for threads in (1, cores):
        t1, times1, blob = _median_seconds(lambda: S.ref_compress(p1, 5, 0, threads=threads), runs)
        cfg1[f"threads_{threads}"] = {
            "Mentries_per_s": p1.size / t1 / 1e6, "median_s": t1, "seconds": times1,
            "byte_exact": hashlib.sha256(blob).hexdigest() == rec["sha256"]}
    return {"value": sample / dt / 1e6, "unit": "Mentries/s", "cores": cores, "kind": "reference",
            "cpu_model": cpu_model,
            "sample": f"2^24-entry Fisher-Yates sample of the workload (input clusters {{10,9}}, "
                      f"output {{9,10}}, embedded_input), reference portable_cpu threads={cores}, "
                      f"median of {runs} runs",
            "median_s": dt, "seconds": times,
            "config1_2e20": {"workload": "2^20 uint32 Fisher-Yates, default format, separate",
                             **cfg1}}


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


def other_inputs(d_out, size, stream, steps: int):
    """The single N-cycle (Sattolo) and identity-with-1%-transpositions inputs
    of the same 2^26 workload, each byte-checked against its digest."""
    out = {}
    for name, label in (("cfg2-cycle-2^26", "single_cycle"), ("cfg2-nearid-2^26", "near_identity_1pct")):
        rec = fixture(name)
        perm = B.generate(rec["kind"], rec["n"], rec["seed"], rec["rate"])
        d_perm = torch.from_numpy(perm.view(np.int32)).cuda()
        del perm

        def step():
            B.compress_device(d_perm.data_ptr(), N, 4, d_out.data_ptr(), size, FMT, stream)

        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        digest = hashlib.sha256(d_out.cpu().numpy().tobytes()).hexdigest()
        out[label] = {"seed": rec["seed"], "rate_ppm": rec["rate"], "ms_per_step": dt * 1e3,
                      "Mentries_per_s": N / dt / 1e6, "byte_exact": digest == rec["sha256"]}
        del d_perm
    return out


def structured_inputs(stream, steps: int):
    """Structured 2^26-scale inputs whose level graphs are long chains for the
    union-find (a rotation by one, an odd-length reverse and perfect shuffle),
    each byte-checked against the reference library's digest."""
    import support as S

    with open(os.path.join(ROOT, "tests", "golden", "structured.json")) as f:
        records = [r for r in json.load(f) if r["n"] >= N]
    out = {}
    for rec in records:
        n = rec["n"]
        fmt = B.FormatOptions(rec["g"], tuple(rec["ic"]), tuple(rec["oc"]), rec["placement"])
        size = B.encoded_size(n, fmt)
        perm = S.structured(rec["kind"], n).astype(np.uint32)
        d_perm = torch.from_numpy(perm.view(np.int32)).cuda()
        del perm
        d_out = torch.empty(size, dtype=torch.uint8, device="cuda")

        def step():
            B.compress_device(d_perm.data_ptr(), n, 4, d_out.data_ptr(), size, fmt, stream)

        step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        digest = hashlib.sha256(d_out.cpu().numpy().tobytes()).hexdigest()
        out[rec["name"]] = {"ms_per_step": dt * 1e3, "Mentries_per_s": n / dt / 1e6,
                            "byte_exact": digest == rec["sha256"]}
        del d_perm, d_out
    return out


def device_lookup(d_out, size: int, d_perm, reps: int = 3):
    """Whole decode of the 2^26 blob on the device (the reference's
    forward_lookup for every entry), checked against the input."""
    d_dec = torch.empty(N, dtype=torch.int32, device="cuda")
    B.lookup_device(d_out.data_ptr(), size, None, N, d_dec.data_ptr(), 4)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        B.lookup_device(d_out.data_ptr(), size, None, N, d_dec.data_ptr(), 4)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    dt = min(times)
    ok = bool(torch.equal(d_dec, d_perm))
    del d_dec
    return {"lookups": N, "ms": dt * 1e3, "Mlookups_per_s": N / dt / 1e6, "round_trip_exact": ok}


def host_inclusive(perm: np.ndarray, want: str | None, reps: int = 5):
    """benes_hip_compress from a host span to a host blob (PCIe both ways):
    the reference's own call shape, reported beside the device-resident rate."""
    B.compress_into(perm, FMT)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        blob = B.compress_into(perm, FMT)   # into a fresh host buffer, as a caller gets one
        times.append(time.perf_counter() - t0)
    dt = sorted(times)[len(times) // 2]
    return {"Mentries_per_s": perm.size / dt / 1e6, "ms": dt * 1e3,
            "device_ms": B.last_device_seconds() * 1e3,
            "byte_exact": want is not None and hashlib.sha256(blob.tobytes()).hexdigest() == want}


def structured(rng, n: int, width: int, near_rate_ppm: int, seed: int) -> np.ndarray:
    u = rng.random()
    kind = B.FISHER_YATES if u < 0.6 else (B.SINGLE_CYCLE if u < 0.8 else B.NEAR_IDENTITY)
    return B.generate(kind, n, seed, near_rate_ppm, width)


def batched(count: int, reps: int = 3, seed: int = 4096):
    """4096 u32 permutations, log-uniform sizes in [2^10, 2^20] (30% shifted
    by U[1, 256]), 60% Fisher-Yates / 20% single cycle / 20% identity with
    0.5% transpositions, middle_group_log2 in {4, 5}, separate placement.
    One batch call against one plain device call per item, inputs resident
    in HBM for both; every batch blob is compared with the plain one."""
    rng = np.random.default_rng(seed)
    perms, fmts = [], []
    for i in range(count):
        n = int(2 ** rng.uniform(10, 20))
        if rng.random() < 0.3:
            n += int(rng.integers(1, 257))
        perms.append(structured(rng, n, 4, 5000, seed * 100000 + i))
        fmts.append(B.FormatOptions(int(rng.integers(4, 6))))
    ns = [p.size for p in perms]
    total = sum(ns)
    offs, pos = [], 0
    for p in perms:
        offs.append(pos)
        pos += (p.nbytes + 7) // 8 * 8
    host = np.zeros(pos, np.uint8)
    for p, o in zip(perms, offs):
        host[o:o + p.nbytes] = p.view(np.uint8)
    del perms
    d_in = torch.from_numpy(host).cuda()
    del host
    ptrs = [d_in.data_ptr() + o for o in offs]
    widths = [4] * count
    out_offs, sizes, need = B.batch_layout(ns, widths, fmts)
    d_batch = torch.zeros(need, dtype=torch.uint8, device="cuda")
    d_plain = torch.zeros(need, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream

    def run_batch():
        return B.compress_batch_device(ptrs, ns, widths, fmts, d_batch.data_ptr(), need, stream)

    def run_plain():
        for p, n, f, o, sz in zip(ptrs, ns, fmts, out_offs, sizes):
            B.compress_device(p, n, 4, d_plain.data_ptr() + o, sz, f, stream)

    # The tuner's trial calls for the batch's size class, and for each plain
    # item's class (every class holds well over eight items), are not timed.
    for _ in range(B.AUTOTUNE_TRIAL_CALLS):
        run_batch()
    run_plain()
    torch.cuda.synchronize()
    tb, tp = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        _, _, errs, _ = run_batch()
        torch.cuda.synchronize()
        tb.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        run_plain()
        torch.cuda.synchronize()
        tp.append(time.perf_counter() - t0)
    tb, tp = min(tb), min(tp)
    same = bool(torch.equal(d_batch, d_plain)) and all(e is None for e in errs)
    return {"workload": f"{count} u32 permutations, log-uniform 2^10..2^20 (30% +U[1,256]), "
                        "60/20/20 Fisher-Yates/single-cycle/near-identity 0.5%, g in {4,5}, separate",
            "items": count, "entries": total, "seed": seed,
            "batch_ms": tb * 1e3, "batch_Mentries_per_s": total / tb / 1e6,
            "plain_ms": tp * 1e3, "plain_Mentries_per_s": total / tp / 1e6,
            "speedup_vs_plain": tp / tb, "batch_equals_plain_bytes": same}


def random_format(rng, n: int) -> "B.FormatOptions":
    g = int(rng.integers(4, 6))
    d = B.layout_of(n, B.FormatOptions(g))["routing_depth"]

    def split():
        if d == 0:
            return ()
        parts = int(rng.integers(1, min(3, d) + 1))
        cuts = sorted(rng.choice(np.arange(1, d), size=parts - 1, replace=False)) if parts > 1 else []
        bounds = [0, *cuts, d]
        return tuple(int(bounds[i + 1] - bounds[i]) for i in range(parts))

    return B.FormatOptions(g, split(), split(), int(rng.integers(0, 3)))


def served(count: int, rate_rps: float, seed: int = 5005, lanes: int = 4, max_batch: int = 512):
    """Requests of log-uniform sizes in [2^8, 2^24], 80% u32 / 20% u64, the
    batched structure mix, random valid cluster splits and placements; host
    arrays in, host blobs out.  Poisson arrivals at `rate_rps` into the
    request loop (64 GB budget) give p50/p99 latency from submit to blob; a
    burst of all requests gives its peak rate; one plain benes_hip_compress
    per request, one after another, is the comparison.  Both paths are timed
    after one untimed pass over the requests (the tuner's trial calls, the
    workspaces' growth), and the server is one long-lived instance."""
    import threading

    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(2 ** rng.uniform(8, 24))
        width = 8 if rng.random() < 0.2 else 4
        reqs.append((structured(rng, n, width, 5000, seed * 100000 + i), random_format(rng, n)))
    total = sum(p.size for p, _ in reqs)
    gaps = rng.exponential(1.0 / rate_rps, size=count)

    # Plain: one call per request, sequentially.
    blobs_plain = []
    lat_plain = []
    # A first untimed pass runs the tuner's trial calls of every size class.
    for p, f in reqs:
        B.compress_into(p, f)
    t0 = time.perf_counter()
    for p, f in reqs:
        t1 = time.perf_counter()
        blobs_plain.append(B.compress_into(p, f))
        lat_plain.append(time.perf_counter() - t1)
    t_plain = time.perf_counter() - t0

    def run(srv, arrivals):
        ids = [None] * count
        before = srv.stats()

        def producer():
            start = time.perf_counter()
            due = 0.0
            for i, (p, f) in enumerate(reqs):
                if arrivals is not None:
                    due += arrivals[i]
                    wait = start + due - time.perf_counter()
                    if wait > 0:
                        time.sleep(wait)
                ids[i] = srv.submit(p, f)

        t0 = time.perf_counter()
        th = threading.Thread(target=producer)
        th.start()
        th.join()
        results = [srv.wait(i) for i in ids]
        wall = time.perf_counter() - t0
        stats = srv.stats()
        stats["batches"] -= before["batches"]
        lats = sorted(lat for _, lat in results)
        same = all(not isinstance(b, Exception) and np.array_equal(b, bp)
                   for (b, _), bp in zip(results, blobs_plain))
        return wall, lats, same, stats

    # One long-lived server, as a service runs: untimed bursts size its
    # lanes' workspaces and staging and run the tuner's trial calls, as the
    # plain calls above had theirs before being timed.  A burst forms about
    # 50 batches over a dozen size classes, some with fewer than the eight
    # trial calls a class takes, so bursts repeat (at most four) until no
    # class is still trying shapes.
    with B.Server(budget_bytes=64 << 30, max_batch=max_batch, lanes=lanes) as srv:
        for _ in range(4):
            run(srv, None)
            if "(trials" not in B.autotune_report():
                break
        wall_p, lats_p, same_p, st_p = run(srv, gaps)
        wall_b, lats_b, same_b, st_b = run(srv, None)
    lp = sorted(lat_plain)
    q = lambda xs, f: xs[min(len(xs) - 1, int(f * len(xs)))] * 1e3
    # Plain calls per size class (log2 of the entries, in steps of 4): the
    # largest class holds most entries, and its plain rate bounds what any
    # batching can gain on this mix.
    classes = {}
    for (p, _), lat in zip(reqs, lat_plain):
        c = min(20, int(np.log2(p.size)) // 4 * 4)
        a = classes.setdefault(f"2^{c}..2^{c + 4}", {"requests": 0, "entries": 0, "seconds": 0.0})
        a["requests"] += 1
        a["entries"] += p.size
        a["seconds"] += lat
    for a in classes.values():
        a["Mentries_per_s"] = a["entries"] / a["seconds"] / 1e6
        a["share_of_entries"] = a["entries"] / total
    return {"workload": f"{count} requests, log-uniform 2^8..2^24, 80% u32 / 20% u64, "
                        "60/20/20 structures, random cluster splits and placements",
            "requests": count, "entries": total, "seed": seed, "lanes": lanes,
            "max_batch": max_batch, "hbm_budget_bytes": 64 << 30,
            "poisson": {"offered_rps": rate_rps, "wall_s": wall_p,
                        "Mentries_per_s": total / wall_p / 1e6, "p50_ms": q(lats_p, 0.5),
                        "p99_ms": q(lats_p, 0.99), "batches": st_p["batches"],
                        "max_concurrent_batches": st_p["max_concurrent_batches"],
                        "peak_bytes": st_p["peak_bytes"]},
            "burst": {"wall_s": wall_b, "Mentries_per_s": total / wall_b / 1e6,
                      "p50_ms": q(lats_b, 0.5), "p99_ms": q(lats_b, 0.99),
                      "batches": st_b["batches"]},
            "plain": {"wall_s": t_plain, "Mentries_per_s": total / t_plain / 1e6,
                      "p50_ms": q(lp, 0.5), "p99_ms": q(lp, 0.99),
                      "by_size_class": dict(sorted(classes.items()))},
            "burst_speedup_vs_plain": t_plain / wall_b,
            "each_result_equals_plain": same_p and same_b}


LARGE_SHAPE = (13, 1024)


def large_u64(seed: int = 0x5EED0000 + 32, reps: int = 2):
    """N = 2^32 + 12345 uint64 entries (34 GB resident in HBM), middle_group_log2
    5, default clusters {14,14}/{14,14}, ranks embedded in the output section.
    The input is a seeded uniformly random permutation drawn on the host's
    threads (Rao-Sandelius: each index to one of 2^14 buckets by its own draw,
    each bucket shuffled by Fisher-Yates; every arrangement is as likely as
    under one sequential Fisher-Yates pass, which at 4.3e9 entries would take
    minutes).  tests/test_gpu_large.py checks this blob byte for byte against
    the independent ruler colouring."""
    n = (1 << 32) + 12345
    fmt = B.FormatOptions(5, (), (), B.EMBEDDED_OUTPUT)
    size = B.encoded_size(n, fmt)
    t0 = time.perf_counter()
    host = B.generate(B.UNIFORM_PARALLEL, n, seed, 0, 8)
    gen_s = time.perf_counter() - t0
    d_perm = torch.from_numpy(host.view(np.int64)).cuda()
    del host
    d_out = torch.empty((size + 7) // 8 * 8, dtype=torch.uint8, device="cuda")
    # Eight tuner trials of 4-7 s each would not fit a bench run: the subtree
    # shape is pinned for this leg (2^13 x 1024: 3.97 s per compress against
    # 4.15 s for 2^14 x 1024, whose LDS holds it to 4 waves per SIMD).
    B.set_subtree_shape(*LARGE_SHAPE)
    times = []
    try:
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            B.compress_device(d_perm.data_ptr(), n, 8, d_out.data_ptr(), d_out.numel(), fmt)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
    finally:
        B.set_subtree_shape(0, 0)
    dt = min(times)
    del d_perm, d_out
    torch.cuda.empty_cache()
    B.release_workspace()
    return {"workload": "plain compress, uniform random permutation of 2^32+12345 uint64 "
                        "(seeded, drawn by Rao-Sandelius bucketing + per-bucket Fisher-Yates), "
                        "g=5, default clusters, embedded_output",
            "n": n, "seed": seed, "blob_bytes": size,
            "ms": dt * 1e3, "Mentries_per_s": n / dt / 1e6, "generate_s": gen_s,
            "reps": reps, "subtree_shape": "%dx%d" % LARGE_SHAPE,
            "checked_by": "tests/test_gpu_large.py (byte-identical to the independent ruler "
                          "colouring; closed forms; round trips)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true",
                    help="only the 2^26 Fisher-Yates line (no other inputs, batch or serving runs)")
    ap.add_argument("--batch-items", type=int, default=4096)
    ap.add_argument("--requests", type=int, default=300)
    ap.add_argument("--offered-rps", type=float, default=400.0)
    ap.add_argument("--no-large", action="store_true", help="skip the 2^32+12345 u64 run")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy (HBM counter calibration)")
    ap.add_argument("--large-only", action="store_true",
                    help="only the 2^32+12345 u64 leg, printed as its own JSON line (profiling)")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched leg, printed as its own JSON line (profiling)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: this backend compresses one permutation (or one "
              "batch) per device and has no multi-GPU path; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    torch.cuda.init()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.large_only:
        assert B.available(), "gfx950 backend unavailable"
        print(json.dumps({"u64_2e32": large_u64()}), flush=True)
        return
    if args.batched_only:
        assert B.available(), "gfx950 backend unavailable"
        print(json.dumps({"batched": batched(args.batch_items)}), flush=True)
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

    # The library tries each subtree shape once on the first calls of a size
    # class and then keeps the fastest; those trial calls are not timed.
    for _ in range(B.AUTOTUNE_TRIAL_CALLS):
        step()
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
    all_stats = B.kernel_stats()
    stats = {k: v for k, v in all_stats.items() if k.startswith("k_")}
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
        "metric": "Mentries/s compressed (byte-exact) at 2^26 u32 + batched/served vs plain; p50/p99",
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
        # Union relations overwritten by a concurrent plain-store hook and
        # joined again by k_union_rejoin, per compress.
        "union_rejoined_per_step": all_stats.get("union_rejoined", (0, 0))[1] / prof_steps,
        # The union pass of every HBM depth together (compare-and-swap
        # launches, plain-store launches, their checks and re-joins): per depth, against
        # its algorithmic bytes (the inverse read, 4 B/entry).
        "union_pass": union_pass(stats, prof_steps),
        "pipeline_roofline": {
            "algorithmic_bytes": 4 * N + 8 * N * 11 + size,
            "kernel_ms": total_ms / prof_steps,
            "frac_of_8TBs": (4 * N + 8 * N * 11 + size) / (ms_per_step * 1e-3) / 8e12,
        },
        "setup": {"generate_s": gen_s},
    }
    # The same compress from a host span to a host blob (PCIe both ways): the
    # reference's own call shape and timing definition, beside the
    # HBM-resident value.
    line["host_inclusive"] = host_inclusive(perm, want)
    if not args.no_extra:
        line["lookup_2e26"] = device_lookup(d_out, size, d_perm)
        line["inputs_2e26"] = other_inputs(d_out, size, stream, max(3, args.steps // 2))
        line["structured_2e26"] = structured_inputs(stream, 3)
        line["batched"] = batched(args.batch_items)
        line["served"] = served(args.requests, args.offered_rps)
        if not args.no_large:
            line["u64_2e32"] = large_u64()
    line["autotune"] = B.autotune_report().strip().splitlines()
    if not args.no_cpu:
        del perm
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
