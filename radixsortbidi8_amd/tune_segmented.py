"""Sweep of the segmented plan's tile configuration and LDS tier.

  python -m radixsortbidi8_amd.tune_segmented [--total-log2 27 | --stated] [--cases i32,...] [--reps 3] [--out PATH]

For every key width and for injective and non-injective key transforms, times the captured plan of a
65,536-segment batch for every (tile_cfg, small_cfg) pair and writes tuning/segmented_table.json:
{"best": {"<bytes>:<injective>": [tile_cfg, small_cfg]}, "times_ms": {...}, "workloads": {...}}.
The batch's lengths are floor(2^U(4, 20)): with --stated as drawn (about 6.2e9 keys), otherwise
scaled to about 2^total keys. --cases limits the sweep to some key types; their entries replace
those of the existing table and the others are kept. The library's automatic choice (tile_cfg =
small_cfg = -1) is the winner of this sweep.
"""
import argparse
import json
import os
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--total-log2", type=int, default=27)
    ap.add_argument("--stated", action="store_true", help="the lengths as drawn, not scaled to 2^total-log2 keys")
    ap.add_argument("--cases", default="u8,i16,i32,f32,u64,f64")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    import torch

    from . import planner as pl
    off = pl.segmented_batch(65536, None if args.stated else args.total_log2)
    total = int(off[-1])
    workload = f"65536 segments, lengths floor(2^U(4, 20)){'' if args.stated else ' scaled'}, {total} keys"
    gen = torch.Generator(device="cuda").manual_seed(0x5EED0007)
    # (name, bytes, injective, dtype of the random bits, dtype sorted, mode)
    cases = (("u8", 1, 1, torch.uint8, torch.uint8, 64), ("i16", 2, 1, torch.int16, torch.int16, 64),
             ("i32", 4, 1, torch.int32, torch.int32, 64), ("f32", 4, 0, torch.int32, torch.float32, 64),
             ("u64", 8, 1, torch.int64, torch.uint64, 64), ("f64", 8, 0, torch.int64, torch.float64, 64))
    wanted = set(args.cases.split(","))
    out = args.out or os.path.join(pl.TUNING_DIR, "segmented_table.json")
    table = {"best": {}, "times_ms": {}, "workloads": {}}
    if os.path.exists(out):
        with open(out) as f:
            table.update(json.load(f))
    table.setdefault("workloads", {})
    old_workload = table.pop("workload", None)
    for name, nb, inj, bits_dt, dt, mode in cases:
        if name not in wanted:
            if old_workload and name not in table["workloads"]:
                table["workloads"][name] = old_workload
            continue
        info = torch.iinfo(bits_dt)
        keys = torch.randint(info.min, info.max, (total,), dtype=bits_dt, device="cuda", generator=gen).view(dt)
        print(f"tune_segmented: {name}, {workload}", flush=True)
        (tc, sc), t = pl.autotune_segmented(keys, off, 0, mode, args.reps,
                                            log=lambda m: print(f"tune_segmented: {name} {m}", flush=True))
        table["best"][f"{nb}:{inj}"] = [tc, sc]
        table["times_ms"][name] = t
        table["workloads"][name] = workload
        print(f"tune_segmented: {name} best tile_cfg {tc} small_cfg {sc} ({t[f'{tc}:{sc}']:.3f} ms)", flush=True)
        del keys
        torch.cuda.empty_cache()
    with open(out, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    print(f"tune_segmented: wrote {out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
