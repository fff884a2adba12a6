"""GPU parity of the keys wider than 64 bits (the 80-bit extended float in 10-, 12- and 16-byte
records, 128-bit integers) against the oracle, through the C ABI: the reference's known answers
(multiplatform/main.cpp:21-178) directly and through device pointers, every mode and direction on
random records with specials and repeats (bit-exact records and permutations), in place and to the
buffer, with a payload, through pointers at both levels, and the allocating forms on host arrays."""
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
r = pytest.importorskip("radixsortbidi8_amd")

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat_wide.json")))
LAYOUT = {"ld80": r.LD80, "ld96": r.LD96, "ld128": r.LD128}
WIDE = [r.LD80, r.LD96, r.LD128, r.U128, r.I128]
MODES = [64, 65, 66] + list(range(8))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def wide_records(keytype, n, rng):
    """n records with random bits, 80-bit specials (zeros of both signs, infinities, NaNs,
    subnormals) or 128-bit specials (0, -1, the minimum), repeats, and random padding."""
    rb, vb = oracle.WIDE_LAYOUT[keytype]
    recs = rng.integers(0, 256, (n, rb), dtype=np.uint8)
    lo = recs[:, :8].copy().view(np.uint64).ravel()
    if vb == 80:
        se = recs[:, 8:10].copy().view(np.uint16).ravel()
        pick = rng.random(n)
        se[pick < 0.05] &= 0x8000                              # zero exponent: zeros and subnormals
        lo[pick < 0.02] = 0
        se[(pick >= 0.05) & (pick < 0.08)] |= 0x7FFF          # infinities and NaNs
        lo[(pick >= 0.05) & (pick < 0.065)] = np.uint64(1 << 63)
        recs[:, 8:10] = se.view(np.uint8).reshape(n, 2)
    else:
        hi = recs[:, 8:16].copy().view(np.uint64).ravel()
        pick = rng.random(n)
        lo[pick < 0.03], hi[pick < 0.03] = 0, 0
        lo[(pick >= 0.03) & (pick < 0.06)] = np.uint64((1 << 64) - 1)
        hi[(pick >= 0.03) & (pick < 0.06)] = np.uint64((1 << 64) - 1)
        lo[(pick >= 0.06) & (pick < 0.08)], hi[(pick >= 0.06) & (pick < 0.08)] = 0, np.uint64(1 << 63)
        recs[:, 8:16] = hi.view(np.uint8).reshape(n, 8)
    recs[:, :8] = lo.view(np.uint8).reshape(n, 8)
    dup = rng.integers(0, n, (n // 8, 2))
    recs[dup[:, 0]] = recs[dup[:, 1]]
    if vb == 80 and rb > 10:
        recs[:, 10:] = rng.integers(0, 256, (n, rb - 10), dtype=np.uint8)  # padding differs between repeats
    return recs


@pytest.mark.parametrize("case", KAT["cases"], ids=lambda c: c["name"])
def test_wide_known_answers_on_gpu(case):
    kt = LAYOUT[case["layout"]]
    rb = oracle.WIDE_LAYOUT[kt][0]
    recs = np.zeros((7, rb), np.uint8)
    for i in range(7):
        recs[i, :8] = np.frombuffer(int(case["input_mantissa"][i]).to_bytes(8, "little"), np.uint8)
        recs[i, 8:] = np.frombuffer(int(case["input_signexponent"][i]).to_bytes(rb - 8, "little"), np.uint8)
    exp = recs[case["expected_perm"]]
    ws = r.Workspace(kt, 7)
    if not case["indirect"]:
        out = torch.empty(7 * rb, dtype=torch.uint8, device="cuda")
        r.radixsortcopynoalloc(dev(recs.ravel()), out, None, ws, case["direction"], case["mode"], keytype=kt)
        assert np.array_equal(host(out).reshape(7, rb), exp)
        h = np.empty_like(recs)
        assert r.radixsortcopy(recs.ravel(), h.ravel(), case["direction"], case["mode"], keytype=kt)
        assert np.array_equal(h, exp)
        return
    d = dev(recs.ravel())
    ptrs = torch.arange(7, dtype=torch.int64, device="cuda") * rb + d.data_ptr()
    o = torch.empty_like(ptrs)
    r.radixsortcopynoalloc_pointers(ptrs, o, ws, kt, 0, case["direction"], case["mode"])
    assert ((host(o) - d.data_ptr()) // rb).tolist() == case["expected_perm"]


@pytest.mark.parametrize("keytype", WIDE)
def test_wide_all_modes_directions(keytype):
    rng = np.random.default_rng(1000 + keytype)
    rb = oracle.WIDE_LAYOUT[keytype][0]
    for n in (2, 3, 1000, 70001):
        recs = wide_records(keytype, n, rng)
        ws = r.Workspace(keytype, n)
        d_in = dev(recs.ravel())
        d_out, d_buf = torch.empty_like(d_in), torch.empty_like(d_in)
        for mode in MODES:
            for d in range(4):
                flags = r.decode(keytype, mode, d)
                perm = oracle.radixsort_wide(keytype, flags, recs)
                r.radixsortcopynoalloc(d_in, d_out, None, ws, d, mode, keytype=keytype)
                assert np.array_equal(host(d_out).reshape(n, rb), recs[perm]), (n, mode, d)
        assert np.array_equal(host(d_in).reshape(n, rb), recs)  # input untouched


@pytest.mark.parametrize("keytype", WIDE)
def test_wide_inplace_pairs_pointers(keytype):
    rng = np.random.default_rng(2000 + keytype)
    rb = oracle.WIDE_LAYOUT[keytype][0]
    n = 300007
    recs = wide_records(keytype, n, rng)
    ws = r.Workspace(keytype, n)
    payload = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    d_rec = dev(recs.ravel())
    order = rng.permutation(n)
    ptrs = torch.from_numpy((d_rec.data_ptr() + order * rb).astype(np.int64)).cuda()
    for mode in (64, 65, 66, 0, 2):
        for d in range(4):
            flags = r.decode(keytype, mode, d)
            perm = oracle.radixsort_wide(keytype, flags, recs)
            for move in (False, True):
                a, b = dev(recs.ravel()), torch.empty_like(d_rec)
                r.radixsortnoalloc(a, b, ws, move, d, mode, keytype=keytype)
                assert np.array_equal(host(b if move else a).reshape(n, rb), recs[perm]), (mode, d, move)
            ko, kb = torch.empty_like(d_rec), torch.empty_like(d_rec)
            po, pb = torch.empty(n, dtype=torch.uint32, device="cuda"), torch.empty(n, dtype=torch.uint32, device="cuda")
            iflags = r.decode(keytype, mode, d, True)
            iperm = oracle.radixsort_wide(keytype, iflags, recs)
            r.sort_pairs(d_rec, ko, kb, dev(payload), po, pb, ws, d, mode, keytype=keytype)
            assert np.array_equal(host(ko).reshape(n, rb), recs[iperm]), (mode, d)
            assert np.array_equal(host(po), payload[iperm]), (mode, d)
            # through device pointers (indirection: the tie reversal is honoured)
            o = torch.empty_like(ptrs)
            r.radixsortcopynoalloc_pointers(ptrs, o, ws, keytype, 0, d, mode)
            pperm = oracle.radixsort_wide(keytype, iflags, recs[order])
            assert np.array_equal((host(o) - d_rec.data_ptr()) // rb, order[pperm]), (mode, d)


def test_wide_second_level_and_allocating_forms():
    """Pointers to 32-byte records holding a pointer to long double values (second level, indexed),
    and the allocating forms on host and device arrays."""
    rng = np.random.default_rng(7)
    n = 20011
    vals = wide_records(r.LD128, 3 * n, rng)            # three long doubles per record
    d_vals = dev(vals.ravel())
    rec = np.zeros((n, 4), np.uint64)
    rec[:, 2] = (d_vals.data_ptr() + np.arange(n, dtype=np.uint64) * 48).astype(np.uint64)
    d_rec = dev(rec)
    ptrs = torch.from_numpy((d_rec.data_ptr() + np.arange(n) * 32).astype(np.int64)).cuda()
    ws = r.Workspace(r.LD128, n)
    o = torch.empty_like(ptrs)
    for d in range(4):
        flags = r.decode(r.LD128, 64, d, True)
        r.radixsortcopynoalloc_pointers(ptrs, o, ws, r.LD128, 16, d, 64, second_level=True, indirection2=0, index2=2)
        exp = oracle.radixsort_wide(r.LD128, flags, vals.reshape(n, 3, 16)[:, 2])
        assert np.array_equal((host(o) - d_rec.data_ptr()) // 32, exp.astype(np.int64)), d
    for kt in (r.LD80, r.I128):
        rb = oracle.WIDE_LAYOUT[kt][0]
        recs = wide_records(kt, 5003, rng)
        for mode, d in ((64, 0), (65, 3), (66, 1)):
            exp = recs[oracle.radixsort_wide(kt, r.decode(kt, mode, d), recs)]
            h = recs.copy()
            assert r.radixsort(h.ravel(), d, mode, keytype=kt)                      # host, in place
            assert np.array_equal(h, exp)
            t = dev(recs.ravel())
            assert r.radixsort(t, d, mode, keytype=kt)                              # device, in place
            assert np.array_equal(host(t).reshape(-1, rb), exp)
