"""GPU parity of the sample partition plan (algorithm 4 for floats, the absolute-value modes and
sorts with a payload: splitters sampled from the keys, two stable partitions, one LDS sort per
bucket) against the oracle, bit for bit, keys and payload, through each of its paths:

* equality buckets: values repeated often enough to take two candidate slots (±0.0, NaN, a few
  integers at 5-20% of the keys) are kept in source order by the partitions alone;
* the per-bucket LDS sort, with the position packed beside (key - min) in one word;
* buckets whose key range leaves no room for the position, or with a run of more than 96 keys of
  one partition value: stable LDS LSD passes (k_bucket_hard);
* buckets above the LDS capacity (8192 keys), forced with a large mean-bucket target: stable LSD
  passes through HBM, in place and into the buffer;
* the presort check, in-place sorts and sorts into the buffer;
* grouping (unstable) partitions for keys-only sorts whose equal sort keys are equal keys, and
  the ranked ones when +0.0 and -0.0 both occur under the float order.
Stability is observed through payloads and through modes where distinct bit patterns rank equal.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
r = pytest.importorskip("radixsortbidi8_amd")

UINT = {4: np.uint32, 8: np.uint64}
FLOAT = {4: np.float32, 8: np.float64}
TORCH = {r.U32: torch.uint32, r.I32: torch.int32, r.F32: torch.float32, r.U64: torch.uint64, r.I64: torch.int64,
         r.F64: torch.float64}


@pytest.fixture
def msd():
    prev = r.set_algorithm(r.ALGO_MSD2)
    yield
    r.set_algorithm(prev)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def heavy_floats(nb, n, rng):
    """N(0,1) floats with ±0.0 (2%), ±inf (1% each), a quiet NaN (1%) and subnormals (1%)."""
    bits = rng.normal(size=n).astype(FLOAT[nb]).view(UINT[nb])
    hi = 8 * nb - 1
    pats = [0, 1 << hi, (0x7F800000 if nb == 4 else 0x7FF0000000000000),
            (0xFF800000 if nb == 4 else 0xFFF0000000000000), (0x7FC00000 if nb == 4 else 0x7FF8000000000000)]
    u = rng.random(n)
    for i, p in enumerate(pats):
        bits[(u >= 0.01 * i) & (u < 0.01 * (i + 1))] = UINT[nb](p)
    sub = (u >= 0.06) & (u < 0.07)
    bits[sub] = rng.integers(1, 1 << 20, int(sub.sum()), dtype=np.uint64).astype(UINT[nb])
    return bits


def heavy_ints(nb, n, rng):
    """Uniform integers, a fifth of them one value, a tenth another, 5% a third, some sign-flipped
    twins (equal under the absolute-value modes)."""
    bits = rng.integers(0, np.iinfo(UINT[nb]).max, n, dtype=UINT[nb], endpoint=True)
    u = rng.random(n)
    v = rng.integers(0, np.iinfo(UINT[nb]).max, 3, dtype=UINT[nb], endpoint=True)
    bits[u < 0.2] = v[0]
    bits[(u >= 0.2) & (u < 0.3)] = v[1]
    bits[(u >= 0.3) & (u < 0.35)] = v[2]
    k = rng.integers(0, n, n // 16)
    h = k.size // 2
    bits[k[:h]] = bits[k[h:2 * h]] ^ UINT[nb](1 << (8 * nb - 1))
    return bits


def sort_pairs(bits, kt, mode, d, payload):
    """sort_pairs on the GPU; payload None: the source index. Returns (keys, payload)."""
    k = dev(bits).view(TORCH[kt])
    ko, kb = torch.empty_like(k), torch.empty_like(k)
    po = torch.empty(bits.size, dtype=torch.uint32, device="cuda")
    pb = torch.empty_like(po)
    ws = r.Workspace(kt, bits.size, r.LAYOUT_PAIRS)
    r.sort_pairs(k, ko, kb, None if payload is None else dev(payload), po, pb, ws, d, mode)
    return host(ko).view(bits.dtype), host(po)


def expected_pairs(bits, kt, mode, d, payload):
    flags = r.decode(kt, mode, d, True)
    if payload is None:
        ek, perm = oracle.radixsort(bits, flags, want_payload=True)
        return ek, perm.astype(np.uint32)
    return oracle.radixsort(bits, flags, payload=payload)


@pytest.mark.parametrize("nb,kt", [(4, r.F32), (8, r.F64), (4, r.I32), (8, r.U64)])
def test_sample_heavy_values_with_payload(nb, kt, msd):
    rng = np.random.default_rng(300 + nb + kt)
    n = 300007
    bits = heavy_floats(nb, n, rng) if kt in (r.F32, r.F64) else heavy_ints(nb, n, rng)
    payload = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    for mode in (64, 65, 66, 1):
        for d in range(4):
            for p in (None, payload):
                got = sort_pairs(bits, kt, mode, d, p)
                exp = expected_pairs(bits, kt, mode, d, p)
                assert np.array_equal(got[0], exp[0]), (mode, d, p is None)
                assert np.array_equal(got[1], exp[1]), (mode, d, p is None)


@pytest.mark.parametrize("nb,kt", [(4, r.F32), (8, r.F64), (4, r.I32), (8, r.I64)])
def test_sample_keys_only_all_directions(nb, kt, msd):
    """Keys-only sorts in every mode, uniform bit patterns and skewed values; the profile shows the
    per-bucket sort ran."""
    rng = np.random.default_rng(400 + nb + kt)
    for n in (1 << 20, 123457):
        for bits in (rng.integers(0, np.iinfo(UINT[nb]).max, n, dtype=UINT[nb], endpoint=True),
                     heavy_floats(nb, n, rng) if kt in (r.F32, r.F64) else heavy_ints(nb, n, rng)):
            a = dev(bits).view(TORCH[kt])
            o, b = torch.empty_like(a), torch.empty_like(a)
            ws = r.Workspace(kt, n)
            for mode in (64, 65, 66):
                for d in range(4):
                    r.profile_begin()
                    r.radixsortcopynoalloc(a, o, b, ws, d, mode)
                    prof = r.profile_end()
                    exp, _ = oracle.radixsort(bits, r.decode(kt, mode, d))
                    assert np.array_equal(host(o).view(bits.dtype), exp), (n, mode, d)
                    if kt in (r.F32, r.F64) or mode != 64:  # (integer native modes: the plan by top bytes)
                        assert prof["msd_local"][1] > 0, (n, mode, d)


def test_sample_wide_ranges_and_long_runs(msd):
    """u32 keys at 2^20: one bucket per first-partition segment, key ranges of 2^24 (no room for the
    position in a 32-bit word), so every bucket takes the LDS LSD passes; then 300 values of about
    870 copies each: runs far above 96 keys in buckets of a narrow range."""
    rng = np.random.default_rng(11)
    n = 1 << 20
    cases = [rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)]
    pool = rng.integers(0, 1 << 32, 300, dtype=np.uint64).astype(np.uint32)
    cases.append(pool[rng.integers(0, 300, n >> 2)])
    for bits in cases:
        payload = rng.integers(0, 1 << 32, bits.size, dtype=np.uint64).astype(np.uint32)
        for d in range(4):
            got = sort_pairs(bits, r.U32, 64, d, payload)
            exp = expected_pairs(bits, r.U32, 64, d, payload)
            assert np.array_equal(got[0], exp[0]) and np.array_equal(got[1], exp[1]), d


@pytest.mark.parametrize("nb,kt,mode", [(4, r.U32, 64), (8, r.F64, 64), (4, r.F32, 65), (8, r.I64, 65)])
def test_sample_buckets_through_hbm(nb, kt, mode, msd):
    """A mean-bucket target of 2^20 keys leaves one bucket per first-partition segment, about 16K
    keys each at 2^22: every bucket exceeds the block's 8192 and takes the LSD passes through HBM
    (in place: the other partition array as the second buffer; copying: the destination)."""
    prev = r.set_sample_target(1 << 20)
    try:
        rng = np.random.default_rng(500 + nb + kt)
        n = (1 << 22) + 1234
        bits = heavy_floats(nb, n, rng) if kt in (r.F32, r.F64) else heavy_ints(nb, n, rng)
        payload = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        for d in (0, 3):
            got = sort_pairs(bits, kt, mode, d, payload)
            exp = expected_pairs(bits, kt, mode, d, payload)
            assert np.array_equal(got[0], exp[0]) and np.array_equal(got[1], exp[1]), d
            exp_k, _ = oracle.radixsort(bits, r.decode(kt, mode, d))
            ws = r.Workspace(kt, n)
            for move in (False, True):
                a = dev(bits).view(TORCH[kt])
                b = torch.empty_like(a)
                r.radixsortnoalloc(a, b, ws, move, d, mode)
                assert np.array_equal(host(b if move else a).view(bits.dtype), exp_k), (d, move)
    finally:
        r.set_sample_target(prev)


@pytest.mark.parametrize("nb,kt", [(4, r.F32), (8, r.F64)])
def test_sample_grouping_partitions_floats(nb, kt, msd):
    """Keys-only float sorts whose equal sort keys are equal keys take grouping (unstable)
    partitions: the one-to-one float modes always, the float order (native) unless both +0.0 and
    -0.0 occur. Zeros of one sign only (grouped), of both signs (ranked: their input order is
    observable and must survive), N(0,1) without zeros, in every direction, bit for bit."""
    rng = np.random.default_rng(600 + nb)
    n = (1 << 20) + 4321
    base = rng.normal(size=n).astype(FLOAT[nb]).view(UINT[nb])
    base[base << 1 == 0] = UINT[nb](1)  # no zeros in the base data
    neg = UINT[nb](1 << (8 * nb - 1))
    u = rng.random(n)
    plus, minus, both = base.copy(), base.copy(), base.copy()
    plus[u < 0.03] = 0
    minus[u < 0.03] = neg
    both[u < 0.02] = 0
    both[(u >= 0.02) & (u < 0.04)] = neg
    a, o, b = None, None, None
    ws = r.Workspace(kt, n)
    for name, bits in (("normal", base), ("+0", plus), ("-0", minus), ("+-0", both)):
        a = dev(bits).view(TORCH[kt])
        o, b = torch.empty_like(a), torch.empty_like(a)
        for mode in (64, r.sortingmode.forcefloatingp, r.sortingmode.specialfloatingp,
                     r.sortingmode.forcetieredabsfloatingp):
            for d in range(4):
                r.radixsortcopynoalloc(a, o, b, ws, d, mode)
                exp, _ = oracle.radixsort(bits, r.decode(kt, mode, d))
                assert np.array_equal(host(o).view(bits.dtype), exp), (name, int(mode), d)


@pytest.mark.parametrize("nb,kt", [(4, r.I32), (8, r.I64)])
def test_sample_grouping_partitions_tiered_abs(nb, kt, msd):
    """Integer keys-only sorts in the tiered absolute-value mode (one to one, so grouped) with
    heavy values, in place and copying."""
    rng = np.random.default_rng(700 + nb)
    n = (1 << 20) + 99
    bits = heavy_ints(nb, n, rng)
    ws = r.Workspace(kt, n)
    for d in range(4):
        exp, _ = oracle.radixsort(bits, r.decode(kt, r.sortingmode.forcetieredabssigned, d))
        a = dev(bits).view(TORCH[kt])
        o, b = torch.empty_like(a), torch.empty_like(a)
        r.radixsortcopynoalloc(a, o, b, ws, d, r.sortingmode.forcetieredabssigned)
        assert np.array_equal(host(o).view(bits.dtype), exp), d
        r.radixsortnoalloc(a, b, ws, False, d, r.sortingmode.forcetieredabssigned)
        assert np.array_equal(host(a).view(bits.dtype), exp), d


def test_sample_2p29_pairs_take_the_large_bucket_blocks():
    """2^29 u64 keys with their source index: the second level's buckets average 8192 keys, so the
    larger ones take the 1024x12 per-bucket blocks; keys and payload checked against a stable
    torch.sort (the payload is the stable permutation)."""
    gen = torch.Generator(device="cuda").manual_seed(729)
    n = 1 << 29
    k = torch.randint(-2**63, 2**63 - 1, (n,), dtype=torch.int64, device="cuda", generator=gen)
    k[: n // 64] = k[n // 2: n // 2 + n // 64]  # duplicates: stability is observable
    ku = k.view(torch.uint64)
    ws = r.Workspace(r.U64, n, r.LAYOUT_PAIRS)
    ko, kb = torch.empty_like(ku), torch.empty_like(ku)
    po = torch.empty(n, dtype=torch.uint32, device="cuda")
    pb = torch.empty_like(po)
    r.profile_begin()
    r.sort_pairs(ku, ko, kb, None, po, pb, ws, 0, 64)
    prof = r.profile_end()
    assert prof["msd_local"][1] > 0
    ref_v, ref_i = torch.sort(k ^ (-(1 << 63)), stable=True)
    assert torch.equal(ko.view(torch.int64) ^ (-(1 << 63)), ref_v)
    assert torch.equal(po.view(torch.int32).to(torch.int64) & 0xFFFFFFFF, ref_i)


def test_sample_target_rejects_negative():
    with pytest.raises(r.Rsbd8Error):
        r.set_sample_target(-1)


def test_sample_presort_check(msd):
    """The presort check of a plan: float keys already in order or strictly reversed become one copy;
    others are sorted."""
    from radixsortbidi8_amd import planner
    rng = np.random.default_rng(21)
    n = 400009
    srt = np.sort(rng.normal(size=n).astype(np.float32))
    ws = r.Workspace(r.F32, n)
    for keys in (srt, srt[::-1].copy(), heavy_floats(4, n, rng).view(np.float32)):
        a = dev(keys)
        o, b = torch.empty_like(a), torch.empty_like(a)
        planner.sort_with_plan(a, o, b, ws, (4, 2, 1), r.decode(r.F32, 64, 0))
        exp, _ = oracle.radixsort(keys.view(np.uint32), r.decode(r.F32, 64, 0))
        assert np.array_equal(host(o).view(np.uint32), exp)


@pytest.mark.parametrize("kt", [r.U32, r.U64, r.I32, r.I64])
def test_algorithm5_takes_integer_keys_with_few_or_skewed_values(kt):
    """Algorithm 5 (the sample partition plan whatever the keys) on keys-only integer sorts that fail
    the exact plan's bucket test: every key equal, 16 distinct values, and a Zipf-like skew (one value
    in about a tenth of the keys), in every direction, with and without the presort check."""
    from radixsortbidi8_amd import planner as pl
    nb = {r.U32: 4, r.U64: 8, r.I32: 4, r.I64: 8}[kt]
    udt = np.uint32 if nb == 4 else np.uint64
    tdt = {r.U32: torch.uint32, r.U64: torch.uint64, r.I32: torch.int32, r.I64: torch.int64}[kt]
    rng = np.random.default_rng(0x5A3 + kt)
    n = (1 << 20) + 7
    ws = r.Workspace(kt, n)
    for shape in ("equal", "unique16", "zipf"):
        if shape == "equal":
            bits = np.full(n, rng.integers(0, np.iinfo(udt).max, dtype=udt), dtype=udt)
        elif shape == "unique16":
            bits = rng.integers(0, np.iinfo(udt).max, 16, dtype=udt, endpoint=True)[rng.integers(0, 16, n)]
        else:
            pool = rng.integers(0, np.iinfo(udt).max, 4096, dtype=udt, endpoint=True)
            bits = pool[np.minimum(rng.zipf(1.3, n) - 1, 4095)]
        a = torch.from_numpy(bits).cuda().view(tdt)
        out, buf = torch.empty_like(a), torch.empty_like(a)
        for d in range(4):
            flags = r.decode(kt, 64, d)
            exp, _ = oracle.radixsort(bits, flags)
            for plan in ((5, -1, 0), (5, -1, 1)):
                pl.sort_with_plan(a, out, buf, ws, plan, flags)
                torch.cuda.synchronize()
                assert np.array_equal(out.cpu().numpy().view(udt), exp), (shape, d, plan)
