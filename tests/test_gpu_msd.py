"""GPU parity of the two-level MSD plan (algorithm 4: partitions by the top two digits, then one
LDS sort per bucket) against the oracle, bit for bit, including the paths a skewed input takes:

* the plan kernel keeps the LSD plan when a top digit is uniform;
* a block whose bucket has a run of more than 96 keys sharing the third digit sorts the rest of
  its digits with LDS LSD passes;
* a bucket larger than the LDS capacity switches the whole sort back to the LSD plan after the
  second count (the input is still intact then), for integer keys under an injective transform;
* keys the plan cannot bucket by their top bytes (floats, the absolute-value modes) take the
  sample partition plan instead, whose buckets come from splitters and so stay small.
Stability is observed through the absolute-value modes, where distinct bit patterns rank equal.
The profile kinds tell which plan ran.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
r = pytest.importorskip("radixsortbidi8_amd")

UINT = {4: np.uint32, 8: np.uint64}
TORCH = {(4, r.U32): torch.uint32, (4, r.I32): torch.int32, (4, r.F32): torch.float32,
         (8, r.U64): torch.uint64, (8, r.I64): torch.int64, (8, r.F64): torch.float64}


@pytest.fixture
def msd():
    prev = r.set_algorithm(r.ALGO_MSD2)
    yield
    r.set_algorithm(prev)


def run(bits, kt, mode, d, inplace=None):
    """Sorts bits (a uint array) as key type kt on the GPU; returns (result bits, msd plan used)."""
    nb = bits.dtype.itemsize
    t = TORCH[(nb, kt)]
    a = torch.from_numpy(bits).cuda()
    ws = r.Workspace(kt, bits.size)
    out, buf = torch.empty_like(a), torch.empty_like(a)
    r.profile_begin()
    if inplace is None:
        r.radixsortcopynoalloc(a.view(t), out.view(t), buf.view(t), ws, d, mode)
        res = out
    else:
        r.radixsortnoalloc(a.view(t), buf.view(t), ws, inplace, d, mode)
        res = buf if inplace else a
    prof = r.profile_end()
    torch.cuda.synchronize()
    return res.cpu().numpy(), prof["msd_local"][1] > 0


def spread(nb, n, rng):
    """Uniformly random keys: every top-two-digit bucket gets about n / 65536 keys."""
    return rng.integers(0, np.iinfo(UINT[nb]).max, n, dtype=UINT[nb], endpoint=True)


@pytest.mark.parametrize("nb,kt", [(4, r.U32), (4, r.I32), (4, r.F32), (8, r.U64), (8, r.I64), (8, r.F64)])
def test_msd_all_directions_and_modes(nb, kt, msd):
    rng = np.random.default_rng(nb * 10 + kt)
    for n in ((1 << 20) + 12345, 300001):
        bits = spread(nb, n, rng)
        # sign-flipped copies of some keys: equal under the absolute-value modes (stability)
        k = rng.integers(0, n, n // 8)
        h = k.size // 2
        bits[k[:h]] = bits[k[h:2 * h]] ^ UINT[nb](1 << (8 * nb - 1))
        for mode in (64, 65, 66):
            for d in range(4):
                got, used = run(bits, kt, mode, d)
                exp, _ = oracle.radixsort(bits, r.decode(kt, mode, d))
                assert np.array_equal(got, exp), (n, mode, d)
                assert used, (n, mode, d)


@pytest.mark.parametrize("nb", [4, 8])
def test_msd_long_runs_take_the_block_lsd_path(nb, msd):
    """A few buckets hold 300-700 keys whose third digit is one value (runs far above 96), with
    abs-equal pairs; they must still fit the LDS capacity chosen for n."""
    rng = np.random.default_rng(41 + nb)
    n = 1 << 21  # mean bucket 32: capacity 1024
    bits = spread(nb, n, rng)
    sh = 8 * nb
    for b, cnt in ((0x0307, 300), (0xF0F0, 700), (0x8001, 500)):
        idx = rng.choice(n, cnt, replace=False)
        low = rng.integers(0, 1 << (sh - 24), cnt, dtype=np.uint64)
        v = (np.uint64(b) << np.uint64(sh - 16)) | (np.uint64(0x5A) << np.uint64(sh - 24)) | low
        v[: cnt // 4] = v[cnt // 4: 2 * (cnt // 4)]  # duplicates
        bits[idx] = v.astype(UINT[nb])
    kt = r.I32 if nb == 4 else r.I64
    for mode in (64, 65):
        for d in range(4):
            got, used = run(bits, kt, mode, d)
            exp, _ = oracle.radixsort(bits, r.decode(kt, mode, d))
            assert np.array_equal(got, exp), (mode, d)
            assert used


@pytest.mark.parametrize("nb,kt,mode", [(4, r.U32, 64), (8, r.U64, 64), (4, r.F32, 64), (8, r.F64, 65)])
def test_msd_oversized_bucket_falls_back(nb, kt, mode, msd):
    """One top-16-bit bucket of 3000 keys with capacity 1024: integer keys under an injective
    transform switch back to the LSD plan (from the intact input); float keys take the sample
    partition plan, whose splitters divide that bucket. ±0.0 among the float keys makes tie order
    observable."""
    rng = np.random.default_rng(77 + nb)
    n = 1 << 20
    bits = spread(nb, n, rng)
    sh = 8 * nb
    idx = rng.choice(n, 3000, replace=False)
    bits[idx] = ((np.uint64(0x4242) << np.uint64(sh - 16)) |
                 rng.integers(0, 1 << (sh - 16), 3000, dtype=np.uint64)).astype(UINT[nb])
    zeros = rng.choice(n, 2000, replace=False)
    bits[zeros[:1000]] = 0
    bits[zeros[1000:]] = UINT[nb](1 << (sh - 1))
    for d in range(4):
        got, used = run(bits, kt, mode, d)
        exp, _ = oracle.radixsort(bits, r.decode(kt, mode, d))
        assert np.array_equal(got, exp), d
        assert used == (kt in (r.F32, r.F64))


@pytest.mark.parametrize("nb", [4, 8])
def test_msd_uniform_top_digit_keeps_lsd(nb, msd):
    """Integer keys whose top digit is one value: the plan by the top two bytes is not taken (n is
    large enough for its exact bucket sizes at both key widths)."""
    rng = np.random.default_rng(5 + nb)
    n = 1 << 20
    bits = spread(nb, n, rng)
    bits &= UINT[nb]((1 << (8 * nb - 8)) - 1)
    bits |= UINT[nb](0x7F << (8 * nb - 8))
    kt = r.U32 if nb == 4 else r.U64
    got, used = run(bits, kt, 64, 0)
    assert np.array_equal(got, np.sort(bits))
    assert not used


@pytest.mark.parametrize("nb", [4, 8])
def test_msd_inplace_and_movetobuffer(nb, msd):
    rng = np.random.default_rng(9 + nb)
    n = (1 << 20) + 7
    bits = spread(nb, n, rng)
    kt = r.I32 if nb == 4 else r.F64
    for mode in (64, 65):
        for d in (0, 3):
            exp, _ = oracle.radixsort(bits, r.decode(kt, mode, d))
            for move in (False, True):
                got, used = run(bits, kt, mode, d, inplace=move)
                assert np.array_equal(got, exp), (mode, d, move)
                assert used


def test_msd_default_plan_at_2p26():
    """The automatic choice at 2^26 u64 keys is the MSD plan; checked against the oracle's
    multithreaded sort (same algorithm as the single-threaded one)."""
    rng = np.random.default_rng(2026)
    n = (1 << 26) + 3
    bits = spread(8, n, rng)
    got, used = run(bits, r.U64, 64, 0)
    exp = np.empty_like(bits)
    oracle.radixsort_mt(bits, 0, 8, exp)
    assert used
    assert np.array_equal(got, exp)


def test_msd_2p29_takes_the_12288_key_tier():
    """2^29 uniform u64 and u32 keys: 8192 keys per bucket on average, so the exact plan needs its
    12288-key tier (1024x12 blocks); checked against torch.sort of the same keys."""
    gen = torch.Generator(device="cuda").manual_seed(529)
    n = 1 << 29
    for dt, it in ((torch.uint64, torch.int64), (torch.uint32, torch.int32)):
        lo, hi = (-2**63, 2**63 - 1) if it == torch.int64 else (-2**31, 2**31 - 1)
        k = torch.randint(lo, hi, (n,), dtype=it, device="cuda", generator=gen).view(dt)
        kt = r.keytype_of(k.dtype)
        ws = r.Workspace(kt, n)
        o, b = torch.empty_like(k), torch.empty_like(k)
        r.profile_begin()
        r.radixsortcopynoalloc(k, o, b, ws, 0, 64)
        prof = r.profile_end()
        assert prof["msd_local"][1] > 0
        sk = k.view(it).to(torch.int64)
        sk = sk ^ (-(1 << 63)) if it == torch.int64 else sk & 0xFFFFFFFF
        ref = torch.sort(sk).values
        got = o.view(it).to(torch.int64)
        got = got ^ (-(1 << 63)) if it == torch.int64 else got & 0xFFFFFFFF
        assert torch.equal(got, ref)
        del k, o, b, ws, sk, ref, got
