"""GPU parity of the segmented (batched) sort: every segment equals the oracle's sort of that
segment alone (the reference sorts each array with its own radixsortcopy call), for short (LDS)
and long (onesweep) segments, empty and one-element segments, all modes, both tie orders, and
through the captured hipGraph replayed several times with eager work in between."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
r = pytest.importorskip("radixsortbidi8_amd")

TORCH_DT = {np.dtype(np.int32): torch.int32, np.dtype(np.uint32): torch.uint32, np.dtype(np.int64): torch.int64,
            np.dtype(np.float32): torch.float32, np.dtype(np.uint16): torch.uint16, np.dtype(np.uint8): torch.uint8,
            np.dtype(np.float64): torch.float64, np.dtype(np.uint64): torch.uint64}


def make_batch(rng, nseg, lo_log2, hi_log2, dtype):
    lens = np.floor(2.0 ** rng.uniform(lo_log2, hi_log2, nseg)).astype(np.int64)
    lens[rng.integers(0, nseg, max(1, nseg // 50))] = 0
    lens[rng.integers(0, nseg, max(1, nseg // 50))] = 1
    off = np.zeros(nseg + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    nb = np.dtype(dtype).itemsize
    bits = rng.integers(0, np.iinfo({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[nb]).max,
                        int(off[-1]), dtype={1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[nb], endpoint=True)
    if nb >= 4:
        bits[::7] &= 0xFF  # duplicates and partly uniform digits
    return off, bits.view(dtype)


def check(keys, off, out, flags):
    exp = oracle.sort_segmented(keys, off, flags)
    if np.array_equal(out.view(np.uint8), exp.view(np.uint8)):
        return
    # name the first segments that differ: length, and whether only the order of equal sort keys
    # (stability) or the sort-key order itself is wrong
    bad = []
    for s in range(off.size - 1):
        a, b = int(off[s]), int(off[s + 1])
        if not np.array_equal(out[a:b].view(np.uint8), exp[a:b].view(np.uint8)):
            ko, ke = oracle.sortkeys(out[a:b], flags), oracle.sortkeys(exp[a:b], flags)
            diff = (out[a:b].view(np.uint8).reshape(b - a, -1) != exp[a:b].view(np.uint8).reshape(b - a, -1)).any(axis=1)
            bad.append((s, b - a, "tie order" if np.array_equal(ko, ke) else "key order", int(np.flatnonzero(diff)[0]),
                        int(diff.sum())))
            if len(bad) >= 6:
                break
    raise AssertionError(f"segments differ (segment, length, kind, first index, differing): {bad}")


@pytest.mark.parametrize("dtype", [np.int32, np.float32, np.uint64, np.uint16, np.uint8])
@pytest.mark.parametrize("small_cfg", [0, 1, 2])
def test_segmented_matches_per_segment_oracle(dtype, small_cfg):
    rng = np.random.default_rng(small_cfg * 10 + np.dtype(dtype).itemsize)
    off, keys = make_batch(rng, 600, 0, 16, dtype)
    d_in = torch.from_numpy(keys).cuda()
    d_out, d_buf = torch.empty_like(d_in), torch.empty_like(d_in)
    kt = r.keytype_of(keys.dtype)
    for mode, d in ((64, 0), (65, 3), (66, 1), (64, 2)):
        plan = r.SegmentedPlan(keys.dtype, off, d, mode, tile_cfg=small_cfg, small_cfg=small_cfg)
        plan.run(d_in, d_out, d_buf)
        torch.cuda.synchronize()
        check(keys, off, d_out.cpu().numpy(), r.decode(kt, mode, d))
        plan.close()


def test_segmented_graph_replays():
    rng = np.random.default_rng(77)
    off, keys = make_batch(rng, 2000, 4, 17, np.int32)
    d_in = torch.from_numpy(keys).cuda()
    d_out, d_buf = torch.empty_like(d_in), torch.empty_like(d_in)
    plan = r.SegmentedPlan(np.int32, off)
    info = plan.info()
    assert info["short"] > 0 and info["long"] > 0
    plan.capture(d_in, d_out, d_buf)
    flags = r.decode(r.I32, 64, 0)
    exp = oracle.sort_segmented(keys, off, flags)
    for rep in range(4):
        d_out.zero_()
        torch.cuda.synchronize()
        plan.launch()
        torch.cuda.synchronize()
        assert np.array_equal(d_out.cpu().numpy(), exp), rep
        # eager work between replays
        r.radixsortcopy(d_in[:1000].clone(), torch.empty(1000, dtype=torch.int32, device="cuda"))
    plan.close()


def test_segmented_all_equal_and_tiny():
    off = np.array([0, 0, 1, 3, 5000, 5000, 20000, 20001], np.uint64)
    keys = np.full(20001, 7, np.int64)
    keys[3:5000] = np.arange(4997) % 3
    d_in = torch.from_numpy(keys).cuda()
    d_out, d_buf = torch.empty_like(d_in), torch.empty_like(d_in)
    for d in range(4):
        plan = r.SegmentedPlan(np.int64, off, d, 64, tile_cfg=0, small_cfg=0)
        plan.run(d_in, d_out, d_buf)
        torch.cuda.synchronize()
        assert np.array_equal(d_out.cpu().numpy(), oracle.sort_segmented(keys, off, r.decode(r.I64, 64, d)))
        plan.close()


@pytest.mark.parametrize("dtype", [np.uint32, np.float64, np.uint16])
def test_segmented_long_equal_runs(dtype):
    """Long segments whose top-digit buckets fit the LDS capacity but hold long runs of keys that
    share the partition bits (random top byte, the rest constant or from a tiny set): the items go
    through the LDS LSD passes; short segments of the same shape take the same route."""
    rng = np.random.default_rng(91 + np.dtype(dtype).itemsize)
    nb = np.dtype(dtype).itemsize
    ut = {2: np.uint16, 4: np.uint32, 8: np.uint64}[nb]
    lens = np.concatenate([rng.integers(20000, 200000, 12), rng.integers(200, 4000, 300)])
    rng.shuffle(lens)
    off = np.zeros(lens.size + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    n = int(off[-1])
    top = rng.integers(0, 256, n, dtype=np.uint64) << np.uint64(8 * nb - 8)
    low = rng.integers(0, 3, n, dtype=np.uint64)
    keys = (top | low).astype(ut).view(dtype)
    d_in = torch.from_numpy(keys).cuda()
    d_out, d_buf = torch.empty_like(d_in), torch.empty_like(d_in)
    kt = r.keytype_of(keys.dtype)
    for small_cfg in (0, 2):
        for mode, d in ((64, 0), (64, 1), (65, 3)):
            plan = r.SegmentedPlan(keys.dtype, off, d, mode, small_cfg=small_cfg)
            assert plan.info()["long"] >= 12
            plan.run(d_in, d_out, d_buf)
            torch.cuda.synchronize()
            check(keys, off, d_out.cpu().numpy(), r.decode(kt, mode, d))
            plan.close()


def test_segmented_autotune_small_batch():
    """The sweep over (tile configuration, LDS tier) runs every pair on a batch and returns one of
    them; the plan it picks sorts the batch like the oracle."""
    from radixsortbidi8_amd import planner as pl
    off = pl.segmented_batch(512, 18, seed=5)
    rng = np.random.default_rng(8)
    keys = rng.integers(-2**31, 2**31, int(off[-1]), dtype=np.int64).astype(np.int32)
    d_in = torch.from_numpy(keys).cuda()
    (tc, sc), times = pl.autotune_segmented(d_in, off, reps=1)
    assert len(times) == len(r.SEGMENTED_TILE_CONFIGS) * len(r.SMALL_CAPS)
    assert f"{tc}:{sc}" in times
    out, buf = torch.empty_like(d_in), torch.empty_like(d_in)
    plan = r.SegmentedPlan(np.int32, off, tile_cfg=tc, small_cfg=sc)
    plan.run(d_in, out, buf)
    torch.cuda.synchronize()
    check(keys, off, out.cpu().numpy(), r.decode(r.I32))
    plan.close()


@pytest.mark.parametrize("dtype", [np.uint32, np.int64, np.float32, np.float64, np.uint16, np.uint8])
@pytest.mark.parametrize("small_cfg", [0, 1, 2])
def test_segmented_tier_boundaries(dtype, small_cfg):
    """Segment lengths at every boundary between the wave kernel (64 keys), the LDS tiers (1024,
    4096, 8192 keys) and the partitioned long segments, in every direction, with duplicates and
    (for floats) signed zeros so tie order is observed."""
    lens = np.array([0, 1, 2, 3, 63, 64, 65, 127, 1023, 1024, 1025, 4095, 4096, 4097, 8191, 8192, 8193, 20000, 70000],
                    np.int64)
    off = np.zeros(lens.size + 1, np.uint64)
    off[1:] = np.cumsum(lens)
    n = int(off[-1])
    nb = np.dtype(dtype).itemsize
    ut = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[nb]
    rng = np.random.default_rng(1000 + 10 * small_cfg + nb)
    bits = rng.integers(0, np.iinfo(ut).max, n, dtype=ut, endpoint=True)
    bits[::5] = bits[::5] & ut(0x0F)  # duplicates
    if np.dtype(dtype).kind == "f":
        z = rng.choice(n, n // 20, replace=False)
        bits[z[: z.size // 2]] = 0
        bits[z[z.size // 2:]] = ut(1 << (8 * nb - 1))
    keys = bits.view(dtype)
    d_in = torch.from_numpy(bits).cuda()
    d_out, d_buf = torch.empty_like(d_in), torch.empty_like(d_in)
    kt = r.keytype_of(keys.dtype)
    for d in range(4):
        plan = r.SegmentedPlan(keys.dtype, off, d, 64, small_cfg=small_cfg)
        plan.run(d_in, d_out, d_buf)
        torch.cuda.synchronize()
        check(keys, off, d_out.cpu().numpy().view(dtype), r.decode(kt, 64, d))
        plan.close()
