"""A short, fixed workload for rocprofv3 runs (kernel traces and counter passes).

  python -m radixsortbidi8_amd.profile_driver [--n-log2 28] [--dtype u64] [--reps 2] [--algo A]
                                               [--tile-config C] [--copy]

Sorts 2^n seeded uniform keys (the benchmark's distribution) reps times after one warm-up sort,
then, with --copy, copies the same number of bytes device to device reps times: a plain copy of
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
    args = ap.parse_args()
    import torch

    import radixsortbidi8_amd as r
    from radixsortbidi8_amd import planner as pl

    n = 1 << args.n_log2
    gen = torch.Generator(device="cuda").manual_seed(0x5EED0001)
    keys = pl.make_keys(args.dtype, n, "uniform", gen)
    out = torch.empty_like(keys)
    buf = torch.empty_like(keys)
    ws = r.Workspace(r.keytype_of(keys.dtype), n)
    r.set_algorithm(args.algo)
    if args.tile_config >= 0:
        r.set_tile_config(args.tile_config)
    r.radixsortcopynoalloc(keys, out, buf, ws)
    torch.cuda.synchronize()
    print(f"profile_driver: warm-up sort of 2^{args.n_log2} {args.dtype} done", flush=True)
    for i in range(args.reps):
        r.radixsortcopynoalloc(keys, out, buf, ws)
        torch.cuda.synchronize()
        print(f"profile_driver: sort {i + 1}/{args.reps} done", flush=True)
    if args.copy:
        for i in range(args.reps):
            buf.copy_(keys)
            torch.cuda.synchronize()
            print(f"profile_driver: copy {i + 1}/{args.reps} of {keys.numel() * keys.element_size()} bytes done",
                  flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
