#!/usr/bin/env python3
"""Headline benchmark: stable LSD radix sort of 2^28 uniform u64 keys on one MI355X.

One step = one complete sort (rsbd8::radixsortcopynoalloc semantics) of the device-resident input
into the output array: the histogram of every digit, the device-side plan, and every executed
partition, scatter and LDS sort launch of the plan it chose.
Prints one JSON line: sorted GKeys/s (value), the HBM roofline of the dominant kernel (of the
kernels that read and write every key, the one with the most time; timed per launch with HIP
events on its stream), the whole-sort roofline fraction with a per-phase breakdown, and the
multithreaded CPU baseline (the oracle's sliced LSD sort) timed on this host.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "sorted GKeys/s & % HBM roofline (u64 2^28); learned vs fixed plan speedup"
HBM_PEAK_GBS = 8000.0


def host_cpus():
    """The host's logical CPU count, the CPUs this process may run on, the thread share the job is
    given (OMP_NUM_THREADS when set, else every CPU it may run on) and the CPU model."""
    nproc = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = nproc
    share = allowed
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        share = min(allowed, int(env))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": nproc, "affinity": allowed, "threads": share, "model": model}


def _multiset_sums(a):
    """Two order-independent sums of a u64 array (plain and mixed, mod 2^64): equal for a
    permutation, and a lost or duplicated key changes them."""
    import numpy as np
    s0 = s1 = 0
    m = np.empty(min(a.size, 1 << 22), np.uint64)   # one reused chunk: no array-sized temporaries
    for i in range(0, a.size, m.size):
        c = a[i:i + m.size]
        t = m[:c.size]
        np.multiply(c, np.uint64(0x9E3779B97F4A7C15), out=t)
        t ^= t >> np.uint64(29)
        s0 += int(c.sum(dtype=np.uint64))
        s1 += int(t.sum(dtype=np.uint64))
    return s0 % (1 << 64), s1 % (1 << 64)


def cpu_baseline(sample_log2: int, budget_s: float):
    """The oracle's LSD sort (the reference's 11-bit digit plan, key range sliced over the job's
    CPU share) on the same workload as the GPU line, verified by order and multiset sums."""
    import numpy as np

    import oracle
    cpus = host_cpus()
    threads = cpus["threads"]
    n = 1 << sample_log2
    rng = np.random.default_rng(0x5EED0001)
    keys = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    out = np.empty_like(keys)
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - t_start < budget_s and len(times) < 7):
        t0 = time.perf_counter()
        oracle.radixsort_mt(keys, 0, threads, out)
        times.append(time.perf_counter() - t0)
    ok = bool((out[1:] >= out[:-1]).all()) and _multiset_sums(out) == _multiset_sums(keys)
    assert ok, "cpu baseline produced a result that is not a sorted permutation"
    med = statistics.median(times)
    return {"value": round(n / med / 1e9, 4), "unit": "GKeys/s", "cores": threads, "kind": "port",
            "host": cpus, "median_s": round(med, 4), "runs": len(times),
            "sample": f"2^{sample_log2} uniform u64 keys (same distribution as the GPU line), oracle LSD sort "
                      f"with the reference's 11-bit digit plan, key range sliced over {threads} threads, "
                      f"median of {len(times)}, checked sorted and a permutation (multiset sums)"}


def learned_vs_fixed(requests: int, seed: int, reps: int):
    """The committed selector (radixsortbidi8_amd/tuning/) against the committed fixed plan table on
    the seeded mixed workload: request_stream(requests, seed, 10, 29), sizes log-uniform in
    2^10-2^29 keys, six dtypes, six distributions. Every request is timed reps times with each plan
    (alternately, from an idle device; the learned one including its selection round trip above
    the size gate); the totals are sums of per-request medians, and every output is compared with
    the fixed plan's. runs: the speed-up of each timing round on its own."""
    from radixsortbidi8_amd import planner as pl
    model = os.path.join(pl.TUNING_DIR, "selector_model.npz")
    if not os.path.exists(model):
        return None
    sel = pl.Selector.load(model)
    table = pl.load_table()

    def log(msg):
        print(f"bench: {msg}", file=sys.stderr, flush=True)

    res = pl.evaluate(pl.request_stream(requests, seed, 10, 29), sel, table, log=log, reps=reps)
    sp = [r["speedup"] for r in res["runs"]]
    return {"requests": res["requests"], "reps": reps, "fixed_ms": round(res["fixed_ms"], 3),
            "learned_ms": round(res["learned_ms"], 3),
            "speedup": round(res["fixed_ms"] / max(res["learned_ms"], 1e-9), 4),
            "speedup_runs": [round(x, 4) for x in sp], "speedup_spread": round(max(sp) - min(sp), 4),
            "mismatches": res["mismatches"], "gate_min_count": sel.min_count,
            "stream": f"planner.request_stream({requests}, {seed:#x}, 10, 29)"}


def traffic_from_profile(n_log2: int, family: str):
    """HBM bytes per launch of the dominant kernel measured with rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by what the same counters read for a plain copy of known size), as recorded
    in profiles/ by the counter runs; None when no such record exists."""
    path = os.path.join(ROOT, "profiles", f"{family}_u64_2p{n_log2}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    return rec.get("bytes_per_launch")


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n-log2", type=int, default=28)
    ap.add_argument("--cpu-sample-log2", type=int, default=28)
    ap.add_argument("--cpu-budget-s", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-selector", action="store_true", help="skip the learned-vs-fixed plan comparison")
    ap.add_argument("--selector-requests", type=int, default=4096,
                    help="requests of the mixed workload (4096: the whole stream the metric is stated on)")
    ap.add_argument("--selector-reps", type=int, default=3, help="timings per request and plan")
    ap.add_argument("--tile-config", type=int, default=-1)
    ap.add_argument("--algo", type=int, default=0, help="0 automatic, 1 onesweep, 2 count + scan, 4 MSD plan")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU (one MI355X); --gpus N with N > 1 is not supported: "
              "the sort runs on one device, and a sort across GPUs would first exchange keys by top-digit ranges",
              file=sys.stderr, flush=True)
        sys.exit(2)

    import torch

    import radixsortbidi8_amd as r

    n = 1 << args.n_log2
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    gen = torch.Generator(device=dev).manual_seed(0x5EED0001)
    keys = torch.randint(-(1 << 63), (1 << 63) - 1, (n,), dtype=torch.int64, device=dev, generator=gen).view(torch.uint64)
    out = torch.empty_like(keys)
    buf = torch.empty_like(keys)
    ws = r.Workspace(r.U64, n)
    if args.tile_config >= 0:
        r.set_tile_config(args.tile_config)
    r.set_algorithm(args.algo)
    stream = torch.cuda.current_stream()

    for _ in range(args.warmup):
        r.radixsortcopynoalloc(keys, out, buf, ws)
    torch.cuda.synchronize()
    print(f"bench: warmup done (n=2^{args.n_log2})", file=sys.stderr, flush=True)

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.radixsortcopynoalloc(keys, out, buf, ws)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    # the same steps again with every launch phase bracketed by HIP events on the launch stream
    # (the events add about 0.1 ms per sort, so the headline steps above run without them)
    r.profile_begin()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for _ in range(args.steps):
        r.radixsortcopynoalloc(keys, out, buf, ws)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    prof = r.profile_end()
    ms_per_step_events = (t3 - t2) * 1e3 / args.steps

    # verify the last result: monotone under the sign-flipped int64 view, and a permutation
    flip = out.view(torch.int64) ^ torch.tensor(-(1 << 63), dtype=torch.int64, device=dev)
    sorted_ok = bool((flip[1:] >= flip[:-1]).all().item())
    ref = torch.sort(keys.view(torch.int64) ^ torch.tensor(-(1 << 63), dtype=torch.int64, device=dev)).values
    perm_ok = bool(torch.equal(ref, flip))
    del ref, flip
    print(f"bench: timed {args.steps} steps, verified sorted={sorted_ok} permutation={perm_ok}", file=sys.stderr,
          flush=True)

    k = 8
    steps = max(1, args.steps)
    # per kernel family: (total ms, launches) over the timed steps, and the algorithmic HBM bytes
    # of one launch (keys read + keys written; the count/histogram reads read every key once)
    fam = {
        "lsd_scatter": (prof["scatter"], 2 * k * n),
        "lsd_count_scan": (prof["countscan"], k * n),
        "msd_scatter": (prof["msd_scatter"], 2 * k * n),
        "msd_local_sort": (prof["msd_local"], 2 * k * n),
        "msd_count_scan": (prof["msd_countscan"], 0),
        "histogram": (prof["hist"], k * n),
    }
    msd = prof["msd_local"][1] > 0
    if msd and prof["msd_countscan"][1] == 0:
        # top-16-bit histogram read, top-digit scatter, segmented second-digit scatter, LDS sort
        sort_bytes = k * n + 2 * k * n + 2 * k * n + 2 * k * n
        passes = 2
        plan = ("exact top-16-bit histogram read, two MSD partitions by the top two 8-bit digits placed by "
                "atomic cursors (per key range, then per bucket of the segment), then one LDS sort per "
                "top-16-bit bucket (an LDS partition on the next 11 bits and exact ranking inside its runs)")
    elif msd:
        # hist read, top-digit scatter, second-digit count read, segmented scatter, LDS sort
        sort_bytes = k * n + 2 * k * n + k * n + 2 * k * n + 2 * k * n
        passes = 2
        plan = ("two MSD partitions by the top two 8-bit digits (count + chained scan + scatter, the second "
                "one segmented), then one LDS sort per top-16-bit bucket (an LDS pass on the third digit and "
                "exact ranking inside its runs)")
    else:
        passes = round(prof["scatter"][1] / steps)
        counted = round(prof["countscan"][1] / steps)
        count_reads = max(0, counted - 1)
        sort_bytes = k * n + count_reads * k * n + passes * 2 * k * n + (2 * k * n if passes == 0 else 0)
        plan = f"8-bit LSD digits, {passes} executed passes, " + (
            "count + chained scan + scatter per pass" if counted else "onesweep scatter per pass")
    # the dominant kernel: of the kernels that read and write every key, the one with the most time
    dom = max(("lsd_scatter", "msd_scatter", "msd_local_sort"), key=lambda f: fam[f][0][0])
    (dom_ms, dom_launches), dom_bytes = fam[dom]
    avg_ms = dom_ms / max(1, dom_launches)
    achieved = dom_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    # the reference's own plan for this key type: one counting read, then six 11-bit-digit passes
    # (H:1848-1896) that each read and write every key
    ref_bytes = k * n + 6 * 2 * k * n
    # the 8-bit LSD plan's bytes (one histogram read, eight read+write passes): the bytes-moved
    # roofline the workload's target is stated against (100% = 136 B/key at 8 TB/s)
    lsd8_bytes = k * n + 8 * 2 * k * n
    sort_gbs = sort_bytes / (ms_per_step * 1e-3) / 1e9
    traffic = traffic_from_profile(args.n_log2, dom)
    phases = {f: {"ms_per_step": round(v[0][0] / steps, 4), "launches_per_step": round(v[0][1] / steps, 2)}
              for f, v in fam.items() if v[0][1]}
    line = {
        "metric": METRIC,
        "value": round(n / (ms_per_step * 1e-3) / 1e9, 4),
        "unit": "GKeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "ms_per_step_with_events": round(ms_per_step_events, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "2^28 uniform u64 keys (seeded), stable ascending sort, keys only, device-resident copy in->out",
            "n": n, "key": "u64", "direction": "ascfwdorder", "mode": "native",
            "plan": plan,
            "tile_config": r.TILE_CONFIGS[args.tile_config] if args.tile_config >= 0 else "auto",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": {"lsd_scatter": "k_scatter (one LSD scatter pass: read + write of every key)",
                       "msd_scatter": "k_scatter (one MSD partition pass: read + write of every key)",
                       "msd_local_sort": "k_local (the per-bucket LDS sort: read + write of every key)"}[dom],
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes_per_launch": dom_bytes,
            "avg_launch_ms": round(avg_ms, 4),
        },
        "sort_roofline": {
            "algorithmic_bytes": sort_bytes,
            "achieved_gbs": round(sort_gbs, 1),
            "frac": round(sort_gbs / HBM_PEAK_GBS, 4),
            "reference_plan_bytes": ref_bytes,
            "reference_plan_frac": round(ref_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "lsd8_plan_bytes": lsd8_bytes,
            "lsd8_plan_frac": round(lsd8_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "phases": phases,
        },
        "verified": sorted_ok and perm_ok,
        "learned_vs_fixed": None,
        "cpu_baseline": None,
    }
    if not args.no_selector:
        print("bench: learned vs fixed plan", file=sys.stderr, flush=True)
        del keys, out, buf, ws
        torch.cuda.empty_cache()
        line["learned_vs_fixed"] = learned_vs_fixed(args.selector_requests, 0x5EED0006, args.selector_reps)
    if not args.no_cpu_baseline:
        print("bench: timing the CPU baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(args.cpu_sample_log2, args.cpu_budget_s)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
