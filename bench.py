"""Benchmark of the SISI intersect-path (merge path) kernels on one MI355X.

Workload (the configuration the headline metric is quoted on): one pair of
sorted unique uint32 sets, |A| = |B| = 2^28, drawn uniformly from [0, 2^32),
count plus materialised result list.  A step is one full intersection
(merge-path kernel + placement kernel) of the HBM-resident pair; the timed
region covers exactly K steps between two stream events.

value = algorithmic bytes per step / seconds per step, with
bytes = 4(|A|+|B|) + 4|A∩B| + 4 (inputs read once, list written once, count).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
METRIC = "SISI input GB/s vs HBM peak; MISI pairs/sec batched vs plain; exact counts"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(path):
    """HBM bytes per launch recorded from rocprofv3 --pmc runs (or None)."""
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def cpu_baseline(coll, nbytes, min_seconds):
    """std::set_intersection (the oracle's restatement of
    std_set_intersection.cpp:68-73) on the same pair, single thread."""
    import oracle
    a, b = coll.set(0), coll.set(1)
    reps, t0 = 0, time.perf_counter()
    while True:
        oracle.materialize(a, b)
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds or reps >= 50:
            break
    return {"value": round(nbytes * reps / el / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"full 2^28 x 2^28 pair, std::set_intersection materialising the list, {reps} reps, "
                      f"{el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=28, help="log2 |A| = log2 |B| (default 28)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--seed", type=int, default=20261203)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the SISI path shards by value "
              f"ranges across GPUs but this build runs and reports a single MI355X.", file=sys.stderr)
        sys.exit(2)

    import set_intersection_amd as si
    from set_intersection_amd import datasets

    n = 1 << args.log2n
    universe = 1 << 32
    log(f"[bench] generating |A|=|B|=2^{args.log2n} uniform in [0,2^32) ...")
    t0 = time.perf_counter()
    coll = datasets.sisi_pair(n, n, universe, args.seed)
    log(f"[bench] generated in {time.perf_counter() - t0:.1f} s")

    ctx = si.Context(0)
    info = ctx.info()
    log(f"[bench] device {info['name']} CUs={info['multiprocessor_count']}")
    ds = ctx.upload(coll, validate=True)
    cap = n
    out = ctx.buffer(cap * 4)
    cnt = ctx.buffer(8)
    cnt_only = ctx.buffer(8)

    # correctness of the measured path against the oracle
    import oracle
    exp = oracle.materialize(coll.set(0), coll.set(1))
    ds.pair_materialize_async(0, 1, out, cnt)
    got_n = int(cnt.download(np.uint64)[0])
    got = out.download(np.uint32, count=got_n)
    exact = bool(got_n == exp.size and np.array_equal(got, exp))
    ds.pair_count_async(0, 1, cnt_only)
    exact = exact and int(cnt_only.download(np.uint64)[0]) == exp.size
    log(f"[bench] |A∩B| = {got_n} (oracle {exp.size}) exact={exact}")
    del got

    nbytes = 4 * (2 * n) + 4 * exp.size + 4

    # --- timed: count + materialise
    ctx.kernel_timing(True)
    for _ in range(args.warmup):
        ds.pair_materialize_async(0, 1, out, cnt)
    ctx.sync()
    ctx.kernel_stats_reset()
    e0 = ctx.record()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        ds.pair_materialize_async(0, 1, out, cnt)
    e1 = ctx.record()
    ctx.sync()
    wall = time.perf_counter() - w0
    ms = ctx.elapsed_ms(e0, e1)
    mat_ms, mat_n = ctx.kernel_stats("ip_sisi_materialize")
    cmp_ms, cmp_n = ctx.kernel_stats("ip_sisi_compact")
    step_s = ms / 1e3 / args.steps
    value = nbytes / step_s / 1e9
    k_avg = mat_ms / max(mat_n, 1) / 1e3
    kernel_bytes = 4 * (2 * n) + 4 * exp.size  # inputs once + staged list once
    achieved = kernel_bytes / k_avg / 1e9

    # --- count only (the reference ip executable's output)
    for _ in range(args.warmup):
        ds.pair_count_async(0, 1, cnt_only)
    ctx.sync()
    ctx.kernel_stats_reset()
    c0 = ctx.record()
    for _ in range(args.steps):
        ds.pair_count_async(0, 1, cnt_only)
    c1 = ctx.record()
    ctx.sync()
    cms = ctx.elapsed_ms(c0, c1) / args.steps
    ck_ms, ck_n = ctx.kernel_stats("ip_sisi_count")
    count_gbps = (4 * 2 * n + 8) / (cms / 1e3) / 1e9
    log(f"[bench] materialise {ms / args.steps:.3f} ms/step ({value:.0f} GB/s), kernel {k_avg * 1e3:.3f} ms "
        f"({achieved:.0f} GB/s), compact {cmp_ms / max(cmp_n, 1):.3f} ms; count-only {cms:.3f} ms "
        f"({count_gbps:.0f} GB/s)")

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        cpu = cpu_baseline(coll, nbytes, args.cpu_seconds)
        log(f"[bench] CPU {cpu['value']} GB/s")

    traffic = read_traffic(os.path.join(ROOT, "profiles", "ip_sisi_materialize_2^28_pmc.md"))
    rec = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"SISI intersect-path count+materialise, |A|=|B|=2^{args.log2n} sorted unique "
                               f"uint32 uniform in [0,2^32), seed {args.seed}",
                   "intersection": int(exp.size), "bytes_per_step": int(nbytes)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "kernel": "k_sisi_merge<256,16,materialise>", "kernel_ms": round(k_avg * 1e3, 4)},
        "cpu_baseline": cpu,
        "exact": exact,
        "count_only": {"ms": round(cms, 4), "GBps": round(count_gbps, 2),
                       "kernel_ms": round(ck_ms / max(ck_n, 1), 4)},
        "wall_s_timed": round(wall, 4),
    }
    print(json.dumps(rec), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
