"""The five workload configurations at their full sizes (seeds as specified for them), checked on the
GPU through properties that hold at any size, plus a bit-exact oracle comparison on a slice small
enough for the oracle to finish in seconds:

C1 2^20 u32 keys, ascending: bit-exact against the oracle.
C2 2^28 f32 N(0,1) with 1% each of +0, -0, +inf, -inf, quiet NaN (both signs) and subnormals,
   u32 source-index payload: ascending native, and nativeabs descending forward and reverse.
   oracle.is_stable_sort proves each result is exactly the stable sort (permutation, gather,
   order by (sort key, source index) in the mode's directions).
C3 2^27 u32 indices into 32-byte records, u64 key at offset 8 with the top 3 bytes constant and
   Zipf(1.1) ranks over 2^20 values: u64 ascending and i64 dscrevorder, same proof through the
   records' keys and the indices' source positions.
C4 65,536 int32 segments with log-uniform lengths scaled to 2^28 keys in total (the stated
   distribution totals about 6.2e9 keys): per-segment order and per-segment multiset checksums on
   the GPU, the oracle on a sample of segments, and the captured graph against the plain run.
C5 4,096 requests of six key types, sizes log-uniform in [2^10, 2^29], six distributions: every
   request through the autotuned fixed plan and through the learned selector, outputs identical,
   each against torch.sort of the same keys, a sample against the oracle, total times reported.
"""
import math
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
r = pytest.importorskip("radixsortbidi8_amd")
from radixsortbidi8_amd import planner as pl  # noqa: E402


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def test_c1_u32_2p20_bit_exact():
    say("C1: 2^20 u32")
    rng = np.random.default_rng(0x5EED0002)
    keys = rng.integers(0, 2**32, 1 << 20, dtype=np.uint64).astype(np.uint32)
    d = torch.from_numpy(keys).cuda()
    out, buf = torch.empty_like(d), torch.empty_like(d)
    r.radixsortcopynoalloc(d, out, buf, r.Workspace(r.U32, keys.size))
    exp, _ = oracle.radixsort(keys, r.decode(r.U32))
    assert np.array_equal(out.cpu().numpy(), exp)


def c2_keys(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal(n, dtype=np.float32)
    bits = x.view(np.uint32)
    m = n // 100
    pos = rng.integers(0, n, 6 * m)
    bits[pos[0 * m:1 * m]] = 0x00000000
    bits[pos[1 * m:2 * m]] = 0x80000000
    bits[pos[2 * m:3 * m]] = 0x7F800000
    bits[pos[3 * m:4 * m]] = 0xFF800000
    bits[pos[4 * m:5 * m]] = np.where(rng.random(m) < 0.5, np.uint32(0x7FC00000), np.uint32(0xFFC00000))
    sub = rng.integers(1, 1 << 23, m, dtype=np.uint64).astype(np.uint32)
    bits[pos[5 * m:6 * m]] = sub | (rng.integers(0, 2, m, dtype=np.uint64).astype(np.uint32) << np.uint32(31))
    return x


def test_c2_f32_2p28_specials_with_payload():
    n = 1 << 28
    say("C2: generating 2^28 f32 keys")
    x = c2_keys(n, 0x5EED0003)
    d_in = torch.from_numpy(x).cuda()
    k_out, k_buf = torch.empty_like(d_in), torch.empty_like(d_in)
    p_out = torch.empty(n, dtype=torch.int32, device="cuda")
    p_buf = torch.empty_like(p_out)
    ws = r.Workspace(r.F32, n, r.LAYOUT_PAIRS)
    for mode, d in ((64, 0), (65, 1), (65, 3)):
        say(f"C2: mode {mode} direction {d}")
        r.sort_pairs(d_in, k_out, k_buf, None, p_out, p_buf, ws, d, mode)
        torch.cuda.synchronize()
        flags = r.decode(r.F32, mode, d, True)
        assert oracle.is_stable_sort(x, k_out.cpu().numpy(), p_out.cpu().numpy().view(np.uint32), flags), (mode, d)
        # bit-exact on a 2^22 slice
        s = x[: 1 << 22]
        ds = d_in[: 1 << 22]
        r.sort_pairs(ds, k_out[: s.size], k_buf[: s.size], None, p_out[: s.size], p_buf[: s.size], ws, d, mode)
        exp_k, exp_p = oracle.radixsort(s, flags, want_payload=True)
        assert np.array_equal(k_out[: s.size].cpu().numpy().view(np.uint32), exp_k.view(np.uint32)), (mode, d)
        assert np.array_equal(p_out[: s.size].cpu().numpy().view(np.uint32), exp_p), (mode, d)


def splitmix64(x):
    z = (x + np.uint64(0x9E3779B97F4A7C15)).astype(np.uint64)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def test_c3_indices_2p27_zipf_records():
    n, stride = 1 << 27, 32
    say("C3: generating 2^27 records")
    rng = np.random.default_rng(0x5EED0004)
    with np.errstate(over="ignore"):
        vals = (np.uint64(0xA5A5A5) << np.uint64(40)) | (splitmix64(np.arange(1 << 20, dtype=np.uint64)) &
                                                          np.uint64((1 << 40) - 1))
    w = np.arange(1, (1 << 20) + 1, dtype=np.float64) ** -1.1
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    ranks = np.minimum(np.searchsorted(cdf, rng.random(n)), (1 << 20) - 1)
    keys = vals[ranks]
    del ranks
    rec = np.zeros(n * 4, np.uint64)
    rec[1::4] = keys
    idx = rng.permutation(n).astype(np.uint32)
    pos = np.empty(n, np.int64)
    pos[idx] = np.arange(n)
    d_rec = torch.from_numpy(rec.view(np.uint8)).cuda()
    d_idx = torch.from_numpy(idx).cuda()
    out = torch.empty_like(d_idx)
    ws = r.Workspace(r.U64, n, r.LAYOUT_INDIRECT)
    logical = keys[idx]
    for dt, kt, d in ((np.uint64, r.U64, 0), (np.int64, r.I64, 3)):
        say(f"C3: {np.dtype(dt).name} direction {d}")
        r.sort_indices(d_rec, stride, 8, d_idx, out, ws, dt, d, 64)
        o = out.cpu().numpy()
        flags = r.decode(kt, 64, d, True)
        assert oracle.is_stable_sort(logical.view(dt), keys[o].view(dt), pos[o], flags), (dt, d)
        # bit-exact on the first 2^20 indices
        sub = torch.empty(1 << 20, dtype=torch.uint32, device="cuda")
        r.sort_indices(d_rec, stride, 8, d_idx[: 1 << 20], sub, ws, dt, d, 64)
        exp = oracle.sort_indirect(rec.view(np.uint8), stride, 8, idx[: 1 << 20], dt, flags)
        assert np.array_equal(sub.cpu().numpy(), exp), (dt, d)


def c4_offsets(nseg, total, seed):
    rng = np.random.default_rng(seed)
    lens = np.floor(2.0 ** rng.uniform(4, 20, nseg))
    lens = np.maximum(1, np.floor(lens * (total / lens.sum()))).astype(np.int64)
    off = np.zeros(nseg + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    return off, lens


def segment_checksums(x, seg_id, nseg):
    """Three multiset checksums per segment (sums of x, x*x and a multiplicative hash, wrapping)."""
    v = x.to(torch.int64)
    out = []
    for f in (v, v * v, v * -7046029254386353131):
        s = torch.zeros(nseg, dtype=torch.int64, device=x.device)
        s.index_add_(0, seg_id, f)
        out.append(s)
    return out


def test_c4_segmented_65536_arrays_2p28_total():
    nseg = 65536
    off, lens = c4_offsets(nseg, 1 << 28, 0x5EED0005)
    total = int(off[-1])
    say(f"C4: {nseg} segments, {total} keys")
    rng = np.random.default_rng(0x5EED0005 + 1)
    keys = rng.integers(-2**31, 2**31, total, dtype=np.int64).astype(np.int32)
    d_in = torch.from_numpy(keys).cuda()
    out, buf = torch.empty_like(d_in), torch.empty_like(d_in)
    plan = r.SegmentedPlan(np.int32, off)
    plan.run(d_in, out, buf)
    torch.cuda.synchronize()
    lens_t = torch.from_numpy(lens).cuda()
    seg_id = torch.repeat_interleave(torch.arange(nseg, device="cuda"), lens_t)
    first = torch.zeros(total, dtype=torch.bool, device="cuda")
    first[torch.from_numpy(off[:-1].astype(np.int64)).cuda()] = True
    ordered = (out[1:] >= out[:-1]) | first[1:]
    assert bool(ordered.all().item())
    for a, b in zip(segment_checksums(d_in, seg_id, nseg), segment_checksums(out, seg_id, nseg)):
        assert torch.equal(a, b)
    # a sample of segments against the oracle
    host = out.cpu().numpy()
    flags = r.decode(r.I32)
    for s in np.random.default_rng(3).choice(nseg, 64, replace=False):
        a, b = int(off[s]), int(off[s + 1])
        exp, _ = oracle.radixsort(keys[a:b], flags)
        assert np.array_equal(host[a:b], exp), s
    # the captured graph replays the same result
    out2 = torch.empty_like(d_in)
    plan.capture(d_in, out2, buf)
    for _ in range(2):
        out2.zero_()
        plan.launch()
        torch.cuda.synchronize()
        assert torch.equal(out, out2)
    plan.close()


def test_c4_segmented_65536_arrays_stated_lengths():
    """C4 as stated: 65,536 int32 arrays with lengths floor(2^U(4, 20)) (seed 0x5EED0005), about
    6.2e9 keys, generated on the device. Every segment's order and three multiset checksums on the
    GPU (in chunks of segments), 64 sampled segments and the longest bit-exact against the oracle,
    and the captured graph's result equal to the plain run's."""
    nseg = 65536
    off = pl.segmented_batch(nseg, None, 0x5EED0005)
    lens = np.diff(off).astype(np.int64)
    total = int(off[-1])
    say(f"C4 stated: {nseg} segments, {total} keys, longest {int(lens.max())}")
    gen = torch.Generator(device="cuda").manual_seed(0x5EED0005 + 1)
    d_in = torch.randint(-2**31, 2**31 - 1, (total,), dtype=torch.int32, device="cuda", generator=gen)
    out, buf = torch.empty_like(d_in), torch.empty_like(d_in)
    plan = r.SegmentedPlan(np.int32, off)
    say(f"C4 stated: plan {plan.info()}")
    plan.run(d_in, out, buf)
    torch.cuda.synchronize()
    say("C4 stated: sorted; checking segment order and checksums")
    first = torch.zeros(total, dtype=torch.bool, device="cuda")
    first[torch.from_numpy(off[:-1].astype(np.int64)).cuda()] = True
    assert bool(((out[1:] >= out[:-1]) | first[1:]).all().item())
    del first
    # per-segment multiset checksums, over chunks of whole segments of at most 2^28 keys
    lo = 0
    while lo < nseg:
        hi = lo
        while hi < nseg and int(off[hi + 1] - off[lo]) <= (1 << 28):
            hi += 1
        a, b = int(off[lo]), int(off[hi])
        seg_id = torch.repeat_interleave(torch.arange(hi - lo, device="cuda"), torch.from_numpy(lens[lo:hi]).cuda())
        for x, y in zip(segment_checksums(d_in[a:b], seg_id, hi - lo), segment_checksums(out[a:b], seg_id, hi - lo)):
            assert torch.equal(x, y), (lo, hi)
        del seg_id
        lo = hi
    flags = r.decode(r.I32)
    sample = list(np.random.default_rng(3).choice(nseg, 64, replace=False)) + [int(np.argmax(lens))]
    for s in sample:
        a, b = int(off[s]), int(off[s + 1])
        exp, _ = oracle.radixsort(d_in[a:b].cpu().numpy(), flags)
        assert np.array_equal(out[a:b].cpu().numpy(), exp), s
    say("C4 stated: graph replay")
    out2 = torch.empty_like(d_in)
    plan.capture(d_in, out2, buf)
    plan.launch()
    torch.cuda.synchronize()
    assert torch.equal(out, out2)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    stream = torch.cuda.current_stream()
    ev[0].record(stream)
    for _ in range(3):
        plan.launch(stream)
    ev[1].record(stream)
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / 3
    say(f"C4 stated: {total} keys in {ms:.2f} ms per graph launch, {total / ms / 1e6:.1f} GKeys/s")
    plan.close()


def test_c5_selector_stream_4096_requests():
    table = pl.load_table()
    sel = pl.Selector.load(pl.TUNING_DIR + "/selector_model.npz")
    reqs = pl.request_stream(4096, 0x5EED0006, 10, 29)
    max_n = max(q[1] for q in reqs)
    arena = pl.Arena(max_n)
    out2 = torch.empty(max_n * 8, dtype=torch.uint8, device="cuda")
    t_fixed = t_learned = 0.0
    checked = 0
    for i, (dt, n, dist, seed) in enumerate(reqs):
        if i % 512 == 0:
            say(f"C5: request {i}/4096 ({dt}, n={n}, {dist}); fixed {t_fixed:.1f} ms, learned {t_learned:.1f} ms")
        gen = torch.Generator(device="cuda").manual_seed(seed)
        keys = pl.make_keys(dt, n, dist, gen)
        flags = r.decode(r.keytype_of(keys.dtype))
        nb = keys.element_size()
        o1, buf = arena.views(keys)
        o2 = out2[: n * nb].view(keys.dtype)
        pf = pl.fixed_plan(table, nb, n)
        pl.sort_with_plan(keys, o1, buf, arena.ws, pf, flags)
        sel.sort(keys, o2, buf, arena.ws, flags, below=pf)
        # the median of three completion times per request and plan, the plans timed alternately
        tf, tl = [], []
        for _ in range(3):
            tf.append(pl.time_call(lambda: pl.sort_with_plan(keys, o1, buf, arena.ws, pf, flags), 1))
            tl.append(pl.time_call(lambda: sel.sort(keys, o2, buf, arena.ws, flags, below=pf), 1))
        t_fixed += float(np.median(tf))
        t_learned += float(np.median(tl))
        assert torch.equal(o1.view(torch.uint8), o2.view(torch.uint8)), (i, dt, n, dist)
        assert torch.equal(pl._sortable(o1), torch.sort(pl._sortable(keys)).values), (i, dt, n, dist)
        if n <= (1 << 16) and i % 4 == 0:
            exp, _ = oracle.radixsort(keys.cpu().numpy(), flags)
            assert np.array_equal(o1.cpu().numpy(), exp), (i, dt, n, dist)
            checked += 1
        del keys
    say(f"C5: fixed {t_fixed:.1f} ms, learned {t_learned:.1f} ms, speedup {t_fixed / t_learned:.4f}, "
        f"{checked} requests against the oracle")
    assert checked > 50
    assert t_learned < t_fixed, (t_fixed, t_learned)
