"""Trains the plan selector on the GPU and writes its artifacts.

  python -m radixsortbidi8_amd.train_selector --out DIR [--train-requests N] [--eval-requests M]

1. autotune: the fixed plan table (fastest plan per key width and size on uniform keys);
2. collect: features and per-plan timings of a seeded request stream generated and timed by the
   sorter itself (sizes log-uniform in [2^10, 2^29], the six dtypes and six distributions);
3. train: the selector network, on the GPU, with the library's forward/backward/Adam kernels
   (optionally each request's squared errors weighted by a power of its best time, see
   request_weights), and the size gate below which the fixed plan is used (the selection's round
   trip does not pay);
4. evaluate: a held-out stream of other seeds: per-plan timings (the regret of the fixed and the
   learned pick against the best plan), then live sorts with the fixed and the learned plan, the
   learned one including its selection cost, outputs compared.
Writes DIR/autotune_table.json, DIR/selector_model.npz, DIR/selector_train.npz,
DIR/selector_heldout.npz, DIR/report.json.
"""
import argparse
import json
import os
import sys
import time

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--train-requests", type=int, default=360)
    ap.add_argument("--eval-requests", type=int, default=240)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--max-log2", type=float, default=29.0)
    ap.add_argument("--weight-power", type=float, default=0.0,
                    help="each training request's squared errors weigh (its best time / the median best "
                         "time) ** p, clipped to [1/16, 16] (0: all alike; 0.5 measured 1.117x on the 4096-request "
                         "stream against 1.098x unweighted for the same retrain)")
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    import torch

    from . import planner as pl

    def log(msg):
        print(msg, flush=True)

    t0 = time.time()
    table = pl.autotune(log=log)
    with open(os.path.join(args.out, "autotune_table.json"), "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    log(f"autotune done in {time.time() - t0:.1f} s")

    train = pl.request_stream(args.train_requests, 0x5EED1000, 10, args.max_log2)
    X, T, M = pl.collect(train, log=log)
    np.savez(os.path.join(args.out, "selector_train.npz"), X=X, T=T, M=M,
             requests=np.array([f"{d}:{n}:{s}" for d, n, s, _ in train]))
    log(f"collect done in {time.time() - t0:.1f} s")

    sel = pl.Selector(seed=3)
    dev = "cuda"
    W = pl.request_weights(T, M, args.weight_power)
    loss = sel.train(torch.from_numpy(X).to(dev), torch.from_numpy(T).to(dev), torch.from_numpy(M * W[:, None]).to(dev),
                     steps=args.steps, log=log)
    overhead = pl.selection_overhead_ms(sel)
    totals = sel.calibrate_gate(train, X, T, M, table, overhead)
    sel.save(os.path.join(args.out, "selector_model.npz"))
    log(f"train done in {time.time() - t0:.1f} s, loss {loss:.4f}, selection overhead {overhead:.4f} ms, "
        f"gate {sel.min_count} keys")

    held = pl.request_stream(args.eval_requests, 0x5EED2000, 10, args.max_log2)
    HX, HT, HM = pl.collect(held, log=log)
    np.savez(os.path.join(args.out, "selector_heldout.npz"), X=HX, T=HT, M=HM,
             requests=np.array([f"{d}:{n}:{s}" for d, n, s, _ in held]))
    offline = pl.regret(held, HX, HT, HM, sel, table)
    log(f"held-out timings: {json.dumps(offline)}")
    res = pl.evaluate(held, sel, table, log=log)
    res["offline"] = offline
    res["train_loss_log2_us"] = loss
    res["train_requests"] = args.train_requests
    res["weight_power"] = args.weight_power
    res["selection_overhead_ms"] = overhead
    res["gate_min_count"] = sel.min_count
    res["gate_train_totals_ms"] = {str(g): round(v, 4) for g, v in totals.items()}
    with open(os.path.join(args.out, "report.json"), "w") as f:
        json.dump(res, f, indent=1)
    log(json.dumps(res))
    return 0 if res["mismatches"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
