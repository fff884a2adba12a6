"""A short, fixed workload for rocprofv3 runs (kernel traces and counter passes).

  python -m radixsortbidi8_amd.profile_driver [--n-log2 28] [--dtype u64] [--reps 2] [--algo A]
                                               [--tile-config C] [--copy] [--normal] [--payload]
                                               [--mode M] [--direction D] [--dist D] [--presort-check]
                                               [--segmented [--small-cfg S] [--sweep]]

Sorts 2^n seeded uniform keys (the benchmark's distribution; --dist picks another of the request
mix, e.g. sorted or reverse; --presort-check sorts with the plan (algorithm, tile configuration,
presort check on), algorithm 4 unless --algo says otherwise) reps times after one warm-up sort,
or (--segmented) sorts a batch of 65,536 segments with log-uniform lengths totalling 2^n keys (int32
keys, or u64/f64 with --seg-dtype; plan tile configuration --tile-config, LDS tier --small-cfg; --sweep
times every pair), then, with --copy, copies the same number of bytes device to device reps times: a plain copy of
known size that calibrates the HBM byte counters (FETCH_SIZE, WRITE_SIZE) against the bytes it
must move.
"""
import argparse
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n-log2", type=int, default=28)
    ap.add_argument("--dtype", default="u64", choices=("u8", "u16", "u32", "u64", "f32", "f64"))
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--algo", type=int, default=0)
    ap.add_argument("--tile-config", type=int, default=-1)
    ap.add_argument("--copy", action="store_true")
    ap.add_argument("--normal", action="store_true",
                    help="f32/f64 keys from N(0,1) instead of uniform bit patterns")
    ap.add_argument("--payload", action="store_true",
                    help="stable key sort carrying the u32 source index (rsbd8_hip_sort_pairs)")
    ap.add_argument("--mode", type=int, default=64, help="sortingmode (64 native, 65 absolute value, ...)")
    ap.add_argument("--direction", type=int, default=0, help="sortingdirection (0 ascfwdorder ... 3 dscrevorder)")
    ap.add_argument("--dist", default="uniform", help="key distribution of planner.make_keys (uniform, sorted, reverse, ...)")
    ap.add_argument("--presort-check", action="store_true", help="sort with the presort check on (a planned sort)")
    ap.add_argument("--segmented", action="store_true",
                    help="instead: 65,536 int32 segments, log-uniform lengths scaled to 2^n keys in total")
    ap.add_argument("--small-cfg", type=int, default=-1, help="segmented: LDS tier of the plan (-1 automatic)")
    ap.add_argument("--seg-dtype", default="i32", choices=("i32", "u64", "f64"), help="segmented: key type")
    ap.add_argument("--sweep", action="store_true", help="segmented: time every tile configuration and LDS tier")
    args = ap.parse_args()
    import torch

    import radixsortbidi8_amd as r
    from radixsortbidi8_amd import planner as pl

    if args.segmented:
        return segmented(args, torch, r)

    n = 1 << args.n_log2
    gen = torch.Generator(device="cuda").manual_seed(0x5EED0001)
    if args.normal and args.dtype in ("f32", "f64"):
        keys = torch.randn(n, dtype=pl.torch_dtype(args.dtype), device="cuda", generator=gen)
    else:
        keys = pl.make_keys(args.dtype, n, args.dist, gen)
    out = torch.empty_like(keys)
    buf = torch.empty_like(keys)
    kt = r.keytype_of(keys.dtype)
    ws = r.Workspace(kt, n, r.LAYOUT_PAIRS if args.payload else r.LAYOUT_KEYS)
    if args.algo != 5:  # algorithm 5 (the sample partition plan whatever the keys) is a plan, not a setting
        r.set_algorithm(args.algo)
    if args.tile_config >= 0:
        r.set_tile_config(args.tile_config)
    if args.payload:
        p_out = torch.empty(n, dtype=torch.int32, device="cuda")
        p_buf = torch.empty_like(p_out)

        def sort():
            r.sort_pairs(keys, out, buf, None, p_out, p_buf, ws, args.direction, args.mode)
    elif args.presort_check or args.algo == 5:
        plan = (args.algo or 4, args.tile_config, 1 if args.presort_check else 0)
        flags = r.decode(kt, args.mode, args.direction)

        def sort():
            pl.sort_with_plan(keys, out, buf, ws, plan, flags)
    else:
        def sort():
            r.radixsortcopynoalloc(keys, out, buf, ws, args.direction, args.mode)
    what = (f"2^{args.n_log2} {args.dtype}{' N(0,1)' if args.normal else ''}"
            f"{'' if args.dist == 'uniform' else ' ' + args.dist}{' + u32 index payload' if args.payload else ''}"
            f"{' with the presort check' if args.presort_check else ''}{' (algorithm 5)' if args.algo == 5 else ''}")
    sort()
    torch.cuda.synchronize()
    print(f"profile_driver: warm-up sort of {what} (mode {args.mode}, direction {args.direction}) done", flush=True)
    for i in range(args.reps):
        sort()
        torch.cuda.synchronize()
        print(f"profile_driver: sort {i + 1}/{args.reps} done", flush=True)
    if args.copy:
        for i in range(args.reps):
            buf.copy_(keys)
            torch.cuda.synchronize()
            print(f"profile_driver: copy {i + 1}/{args.reps} of {keys.numel() * keys.element_size()} bytes done",
                  flush=True)
    return 0


def segmented(args, torch, r):
    import numpy as np
    rng = np.random.default_rng(0x5EED0005)
    nseg = 65536
    lens = np.floor(2.0 ** rng.uniform(4, 20, nseg))
    lens = np.maximum(1, np.floor(lens * ((1 << args.n_log2) / lens.sum()))).astype(np.int64)
    off = np.zeros(nseg + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    total = int(off[-1])
    gen = torch.Generator(device="cuda").manual_seed(0x5EED0005)
    if args.seg_dtype in ("u64", "f64"):
        keys = torch.randint(-(1 << 63), (1 << 63) - 1, (total,), dtype=torch.int64, device="cuda", generator=gen)
        npdt = np.uint64 if args.seg_dtype == "u64" else np.float64
        keys = keys.view(torch.uint64 if args.seg_dtype == "u64" else torch.float64)
    else:
        keys = torch.randint(-(1 << 31), (1 << 31) - 1, (total,), dtype=torch.int32, device="cuda", generator=gen)
        npdt = np.int32
    out, buf = torch.empty_like(keys), torch.empty_like(keys)
    pairs = [(args.tile_config, args.small_cfg)]
    if args.sweep:
        pairs = [(tc, sc) for sc in range(len(r.SMALL_CAPS)) for tc in range(len(r.SEGMENTED_TILE_CONFIGS))]
    for tc, sc in pairs:
        plan = r.SegmentedPlan(npdt, off, tile_cfg=tc, small_cfg=sc)
        print(f"profile_driver: segmented plan tile_cfg {tc} small_cfg {sc} {plan.info()}", flush=True)
        plan.run(keys, out, buf)
        plan.capture(keys, out, buf)
        torch.cuda.synchronize()
        for name, fn in (("run", lambda: plan.run(keys, out, buf)), ("graph", lambda: plan.launch())):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            for _ in range(args.reps):
                fn()
            b.record()
            b.synchronize()
            ms = a.elapsed_time(b) / args.reps
            print(f"profile_driver: segmented {name}: {ms:.3f} ms per batch, {total / ms / 1e6:.2f} GKeys/s "
                  f"({nseg} segments, {total} keys; tile_cfg {tc} small_cfg {sc})", flush=True)
        plan.close()
    if args.copy:
        for i in range(args.reps):
            buf.copy_(keys)
            torch.cuda.synchronize()
            print(f"profile_driver: copy {i + 1}/{args.reps} of {keys.numel() * keys.element_size()} bytes done",
                  flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
