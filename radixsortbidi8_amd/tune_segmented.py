"""Sweep of the segmented plan's tile configuration and LDS tier.

  python -m radixsortbidi8_amd.tune_segmented [--total-log2 27] [--reps 3] [--out PATH]

For every key width and for injective and non-injective key transforms, times the captured plan of a
65,536-segment batch (log-uniform lengths 2^4..2^20, about 2^total keys) for every (tile_cfg,
small_cfg) pair and writes tuning/segmented_table.json: {"best": {"<bytes>:<injective>": [tile_cfg,
small_cfg]}, "times": {...}}. The library's automatic choice (tile_cfg = small_cfg = -1) is the winner
of this sweep.
"""
import argparse
import json
import os
import sys


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--total-log2", type=int, default=27)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    import torch

    from . import planner as pl
    off = pl.segmented_batch(65536, args.total_log2)
    total = int(off[-1])
    gen = torch.Generator(device="cuda").manual_seed(0x5EED0007)
    # (name, bytes, injective, dtype of the random bits, dtype sorted, mode)
    cases = (("u8", 1, 1, torch.uint8, torch.uint8, 64), ("i16", 2, 1, torch.int16, torch.int16, 64),
             ("i32", 4, 1, torch.int32, torch.int32, 64), ("f32", 4, 0, torch.int32, torch.float32, 64),
             ("u64", 8, 1, torch.int64, torch.uint64, 64), ("f64", 8, 0, torch.int64, torch.float64, 64))
    best, times = {}, {}
    for name, nb, inj, bits_dt, dt, mode in cases:
        info = torch.iinfo(bits_dt)
        keys = torch.randint(info.min, info.max, (total,), dtype=bits_dt, device="cuda", generator=gen).view(dt)
        (tc, sc), t = pl.autotune_segmented(keys, off, 0, mode, args.reps,
                                            log=lambda m: print(f"tune_segmented: {name} {m}", flush=True))
        best[f"{nb}:{inj}"] = [tc, sc]
        times[name] = t
        print(f"tune_segmented: {name} best tile_cfg {tc} small_cfg {sc} ({t[f'{tc}:{sc}']:.3f} ms)", flush=True)
        del keys
    out = args.out or os.path.join(pl.TUNING_DIR, "segmented_table.json")
    with open(out, "w") as f:
        json.dump({"workload": f"65536 segments, log-uniform lengths 2^4..2^20, {total} keys",
                   "best": best, "times_ms": times}, f, indent=1, sort_keys=True)
    print(f"tune_segmented: wrote {out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
