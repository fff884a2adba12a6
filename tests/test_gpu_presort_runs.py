"""GPU parity of the presort check's run reversal: an input that, read in its logical order, never
ascends in sort-key order but has ties is sorted by moving the runs of equal keys as blocks, each
run in its own order (the stable order). Inputs are the oracle's sorted output read backwards
(forward directions) or as is (reverse-reading directions), for every key width, with ties across
tile boundaries (long runs), ±0.0 ties under the float order, in every direction; every result is
checked bit for bit against the oracle. Strictly reversed and unsorted inputs still take their
own paths."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
r = pytest.importorskip("radixsortbidi8_amd")
from radixsortbidi8_amd import planner as pl  # noqa: E402

UINT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
KT = {r.U8: 1, r.U16: 2, r.U32: 4, r.U64: 8, r.I32: 4, r.I64: 8, r.F32: 4, r.F64: 8}
TORCH = {r.U8: torch.uint8, r.U16: torch.uint16, r.U32: torch.uint32, r.U64: torch.uint64, r.I32: torch.int32,
         r.I64: torch.int64, r.F32: torch.float32, r.F64: torch.float64}


def keys_with_ties(kt, n, rng, pool):
    nb = KT[kt]
    vals = rng.integers(0, np.iinfo(UINT[nb]).max, pool, dtype=UINT[nb], endpoint=True)
    bits = vals[rng.integers(0, pool, n)]
    if kt in (r.F32, r.F64):
        # finite values only, and both zero signs (equal under the float order)
        bits &= UINT[nb]((1 << (8 * nb - 2)) - 1)
        u = rng.random(n)
        bits[u < 0.05] = 0
        bits[(u >= 0.05) & (u < 0.1)] = UINT[nb](1 << (8 * nb - 1))
    return bits


def run_plan(bits, kt, flags, plan):
    a = torch.from_numpy(bits).cuda().view(TORCH[kt])
    out, buf = torch.empty_like(a), torch.empty_like(a)
    ws = r.Workspace(kt, bits.size)
    pl.sort_with_plan(a, out, buf, ws, plan, flags)
    torch.cuda.synchronize()
    return out.cpu().numpy().view(bits.dtype)


@pytest.mark.parametrize("kt", [r.U8, r.U16, r.U32, r.U64, r.I32, r.F32, r.F64])
def test_reverse_with_ties_takes_the_run_reversal(kt):
    rng = np.random.default_rng(900 + kt)
    for n in (2, 4097, 3 * 4096 + 5, (1 << 20) + 17):
        for pool in (3, max(2, n // 5)):  # long runs across tiles; short runs
            raw = keys_with_ties(kt, n, rng, pool)
            for mode in (64, 65) if kt not in (r.U8, r.U16) else (64,):
                for d in range(4):
                    flags = r.decode(kt, mode, d)
                    srt, _ = oracle.radixsort(raw, flags)
                    # logical order never ascending: the sorted keys backwards, or as they are
                    # when the sort reads back to front
                    bits = srt if d in (2, 3) else srt[::-1].copy()
                    exp, _ = oracle.radixsort(bits, flags)
                    for plan in ((2, -1, 1), (4, -1, 1)):
                        got = run_plan(bits, kt, flags, plan)
                        assert np.array_equal(got, exp), (n, pool, mode, d, plan)


@pytest.mark.parametrize("kt", [r.U32, r.U64, r.F64])
def test_presort_other_verdicts_unchanged(kt):
    """Strictly reversed input (a reversed copy) and sorted input (a copy), both made by the check
    itself while it reads; unsorted input (a sort); and sorted or strictly reversed inputs with one
    swap early, in the middle or at the very end: the check copies the tiles before it, then a real
    sort overwrites them. The output array keeps the previous sort's keys between calls."""
    rng = np.random.default_rng(950 + kt)
    n = (1 << 18) + 3
    nb = KT[kt]
    raw = rng.permutation(n).astype(UINT[nb])
    if kt == r.F64:
        raw = rng.permutation(n).astype(np.float64).view(np.uint64)
    for d in (0, 2):
        flags = r.decode(kt, 64, d)
        srt, _ = oracle.radixsort(raw, flags)
        phys = srt if d == 0 else srt[::-1].copy()  # in order when read in the sort's direction
        cases = [phys[::-1].copy(), phys, raw]
        for at in (5, n // 2, n - 2):
            for base in (phys, phys[::-1]):  # nearly sorted, nearly strictly reversed
                near = base.copy()
                near[at], near[at + 1] = near[at + 1], near[at]
                cases.append(near)
        cases.append(phys[::-1].copy())  # a reversed copy again after the partial copies
        for plan in ((2, -1, 1), (4, -1, 1), (1, -1, 1)):
            a = torch.empty(n, dtype=TORCH[kt], device="cuda")
            out, buf = torch.empty_like(a), torch.empty_like(a)
            ws = r.Workspace(kt, n)
            for bits in cases:
                exp, _ = oracle.radixsort(bits, flags)
                a.copy_(torch.from_numpy(bits).view(TORCH[kt]))
                pl.sort_with_plan(a, out, buf, ws, plan, flags)
                torch.cuda.synchronize()
                got = out.cpu().numpy().view(bits.dtype)
                assert np.array_equal(got, exp), (d, plan)


@pytest.mark.parametrize("kt", [r.U8, r.U16, r.U32, r.U64])
def test_run_reversal_under_abs_and_tiered_modes(kt):
    """Ties of unequal keys: under the absolute-value modes x and -x share a sort key, under the
    tiered ones the sort key drops the sign bit's place. Runs are runs of equal sort keys."""
    rng = np.random.default_rng(970 + kt)
    for n in (4097, (1 << 16) + 3):
        raw = keys_with_ties(kt, n, rng, 40)
        raw[::3] = (UINT[KT[kt]](0) - raw[::3])  # negations of pool values: ties under the abs modes
        for mode in (1, 3):
            for d in range(4):
                flags = r.decode(kt, mode, d)
                srt, _ = oracle.radixsort(raw, flags)
                bits = srt if d in (2, 3) else srt[::-1].copy()
                exp, _ = oracle.radixsort(bits, flags)
                got = run_plan(bits, kt, flags, (2, -1, 1))
                assert np.array_equal(got, exp), (n, mode, d)


@pytest.mark.parametrize("kt", [r.U8, r.U16])
def test_presort_check_orders_narrow_keys_in_their_width(kt):
    """Under the tiered absolute-value order a 1- or 2-byte key's transform moves the sign bit above
    the key's width. An input ascending when that bit is counted, but not in the real order, must
    be sorted, not copied."""
    nb = KT[kt]
    rng = np.random.default_rng(990 + kt)
    n = 5000
    raw = rng.integers(0, 1 << (8 * nb), n, dtype=np.uint64)
    width = np.uint64((1 << (8 * nb)) - 1)
    neg = raw >> np.uint64(8 * nb - 1)
    # the key with the sign bit above the width
    wide = ((raw << np.uint64(1)) & ((width << np.uint64(1)) | np.uint64(1))) ^ np.where(neg == 1, width, np.uint64(0))
    bits = raw[np.argsort(wide, kind="stable")].astype(UINT[nb])
    flags = r.decode(kt, 1, 0)
    exp, _ = oracle.radixsort(bits, flags)
    assert not np.array_equal(exp, bits)
    got = run_plan(bits, kt, flags, (2, -1, 1))
    assert np.array_equal(got, exp)
