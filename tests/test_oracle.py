"""Pins the CPU oracle (oracle/rsbd8_oracle.c) before it is trusted as the checker.

* Known answers from the reference's own tests and documentation (tests/golden/kat_reference.json,
  multiplatform/main.cpp:10-215 and Radixsortbidi8.hpp:226-249, 265-274, 405-413).
* The key transforms against orderings derived independently from numpy's IEEE floats and
  signed integers (not a retyping of the transform), brute force over every 16-bit pattern.
* The radix sort against an independent merge sort and numpy's stable argsort, plus
  permutation / monotonicity / stability invariants and the degenerate sizes.
"""
import json
import os

import numpy as np
import pytest

import oracle
from oracle import ABS, DESC, FLTP, REV, SIGN

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat_reference.json")))

UINT = {8: np.uint8, 16: np.uint16, 32: np.uint32, 64: np.uint64}
SINT = {8: np.int8, 16: np.int16, 32: np.int32, 64: np.int64}
FLOAT = {16: np.float16, 32: np.float32, 64: np.float64}
MODES = [0, 1, 2, 3, 4, 5, 6, 7]          # explicit sortingmode values, H:224-248
DIRECTIONS = [0, 1, 2, 3]                  # sortingdirection values, H:256-263


def _bits_of(case, key="input"):
    dt = np.dtype(case["dtype"])
    if key + "_bits" in case:
        return np.array(case[key + "_bits"], dtype=UINT[dt.itemsize * 8]).view(dt)
    return np.array(case[key], dtype=dt)


# ---- known answers -------------------------------------------------------------------------

@pytest.mark.parametrize("case", KAT["cases"], ids=lambda c: c["name"])
def test_reference_known_answers(case):
    keys = _bits_of(case)
    flags = oracle.decode(oracle.keytype_of(keys.dtype), case["mode"], case["direction"], case["indirect"])
    for sorter in ("radix", "merge"):
        if sorter == "radix":
            out, perm = oracle.radixsort(keys, flags, want_payload=True)
        else:
            out, perm = oracle.mergesort(keys, flags)
        if "expected_perm" in case:
            assert perm.tolist() == case["expected_perm"], sorter
        if "expected" in case or "expected_bits" in case:
            exp = _bits_of(case, "expected")
            assert out.view(UINT[keys.dtype.itemsize * 8]).tolist() == exp.view(UINT[keys.dtype.itemsize * 8]).tolist(), sorter


@pytest.mark.parametrize("case", KAT["documented_orders"], ids=lambda c: c["name"])
def test_documented_mode_orders(case):
    dt = np.dtype(case["dtype"])
    order = np.array(case["order_bits"], dtype=UINT[dt.itemsize * 8])
    flags = oracle.decode(oracle.keytype_of(dt), case["mode"], 0, False)
    rng = np.random.default_rng(7)
    for _ in range(20):
        shuffled = order[rng.permutation(order.size)]
        out, _ = oracle.radixsort(shuffled.view(dt), flags)
        assert out.view(order.dtype).tolist() == order.tolist()
        # descending is the exact reverse for distinct keys
        out, _ = oracle.radixsort(shuffled.view(dt), flags | DESC)
        assert out.view(order.dtype).tolist() == order[::-1].tolist()


def test_decode_matches_documented_defaults():
    """Mode decoding against the enum documentation (H:212-265), not against the code."""
    named = {"forceunsigned": 0, "forcesigned": SIGN, "forceabssigned": ABS | SIGN,
             "forcefloatingp": SIGN | FLTP, "forceabsfloatingp": ABS | SIGN | FLTP,
             "specialfloatingp": FLTP, "forcetieredabsfloatingp": ABS | FLTP,
             "forcetieredabssigned": ABS}
    enum_value = {"forceunsigned": 0, "forcesigned": 2, "forceabssigned": 3, "forcefloatingp": 6,
                  "forceabsfloatingp": 7, "specialfloatingp": 4, "forcetieredabsfloatingp": 5,
                  "forcetieredabssigned": 1}
    for name, v in enum_value.items():
        for kt in range(10):
            assert oracle.decode(kt, v, 0) == named[name], (name, kt)
    for kt in (oracle.U8, oracle.U16, oracle.U32, oracle.U64):
        # "no distinct effect when used on an unsigned integer input type" (H:218, H:221)
        for native in (64, 65, 66):
            assert oracle.decode(kt, native, 0) == named["forceunsigned"]
    for kt in (oracle.I8, oracle.I16, oracle.I32, oracle.I64):
        assert oracle.decode(kt, 64, 0) == named["forcesigned"]            # H:229-230
        assert oracle.decode(kt, 65, 0) == named["forceabssigned"]         # H:216-217
        assert oracle.decode(kt, 66, 0) == named["forcetieredabssigned"]   # H:219-220
    for kt in (oracle.F32, oracle.F64):
        assert oracle.decode(kt, 64, 0) == named["forcefloatingp"]         # H:233-234
        assert oracle.decode(kt, 65, 0) == named["forceabsfloatingp"]
        assert oracle.decode(kt, 66, 0) == named["forcetieredabsfloatingp"]
    # direction bits (H:253-263); isrevorder is always honoured with indirection
    for d in DIRECTIONS:
        f = oracle.decode(oracle.U32, 64, d, True)
        assert bool(f & DESC) == bool(d & 1) and bool(f & REV) == bool(d & 2)


@pytest.mark.parametrize("mode", MODES)
def test_dropped_revorder_is_invisible(mode):
    """For plain arrays the reference drops isrevorder where equal keys imply equal bits
    (H:59877-59904); check that dropping it cannot change any output."""
    rng = np.random.default_rng(mode)
    keys = rng.integers(0, 1 << 16, 4096, dtype=np.uint16)
    keys[:2000] = keys[2000:4000]  # plenty of duplicates
    f = oracle.decode(oracle.I16, mode, 0, True)
    a, _ = oracle.radixsort(keys, f)
    b, _ = oracle.radixsort(keys, f | REV)
    kept = oracle.decode(oracle.I16, mode, 2, False) & REV
    if not kept:
        assert a.tolist() == b.tolist()


# ---- key transforms against independent orderings ------------------------------------------

def _independent_order(bits, w, mode):
    """Stable order of the patterns implied by the mode's documented semantics, computed with
    numpy integers/floats (lexsort: last key is primary)."""
    u = bits.astype(np.uint64)
    s = (u >> np.uint64(w - 1)).astype(np.int64)
    si = bits.view(SINT[w]).astype(np.int64)
    mag_int = np.where(si < 0, (-(si + 1)).astype(np.uint64) + np.uint64(1), si.astype(np.uint64))
    if mode == 0:
        return np.lexsort((u,))
    if mode == SIGN:
        return np.lexsort((si,))
    if mode == ABS | SIGN:
        return np.lexsort((mag_int,))
    if mode == ABS:
        return np.lexsort((s == 0, mag_int))
    f = bits.view(FLOAT[w]).astype(np.float64)
    nan = np.isnan(f)
    payload = (u & ~(np.uint64(1) << np.uint64(w - 1))).astype(np.float64)
    absf = np.abs(np.where(nan, 0.0, f))
    if mode == SIGN | FLTP:       # IEEE order, -NaN below everything, +NaN above, -0 == +0
        cls = np.where(nan, np.where(s == 1, 0, 2), 1)
        inner = np.where(nan, np.where(s == 1, -payload, payload), 0.0)
        return np.lexsort((inner, np.where(nan, 0.0, f), cls))
    if mode == ABS | SIGN | FLTP:  # |x|, NaNs above inf by payload, x == -x
        return np.lexsort((np.where(nan, payload, 0.0), absf, nan))
    if mode == FLTP:               # +0 .. +inf, +NaN, -NaN, -inf .. -0 (H:241-243)
        neg = s == 1
        sub = np.where(neg, ~nan, nan)
        inner = np.where(neg, np.where(nan, -payload, -absf), np.where(nan, payload, absf))
        return np.lexsort((inner, sub, neg))
    if mode == ABS | FLTP:         # -0, +0, -min, +min, ... -inf, +inf, NaNs (H:244-246)
        return np.lexsort((s == 0, np.where(nan, payload, absf), nan))
    raise AssertionError(mode)


def _patterns(w, rng):
    if w == 16:
        return np.arange(1 << 16, dtype=np.uint64).astype(np.uint16)
    n = 1 << 15
    r = rng.integers(0, np.iinfo(UINT[w]).max, n, dtype=UINT[w], endpoint=True)
    specials = []
    msb = 1 << (w - 1)
    e = {32: 0x7F800000, 64: 0x7FF0000000000000}[w]
    for v in (0, 1, 2, msb - 1, e, e | 1, e | (e >> 9 if w == 32 else e >> 12), msb >> 1):
        specials += [v, v | msb]
    return np.concatenate([np.array(specials, dtype=UINT[w]), r, r[:4096]])  # duplicates too


@pytest.mark.parametrize("w", [16, 32, 64])
@pytest.mark.parametrize("mode", MODES)
def test_sortkey_orders(w, mode):
    rng = np.random.default_rng(w * 16 + mode)
    bits = _patterns(w, rng)
    rng.shuffle(bits)
    flags = oracle.decode(oracle.I8, mode, 0)  # explicit modes ignore the type
    keys, perm = oracle.mergesort(bits, flags)
    exp = _independent_order(bits, w, mode)
    assert np.array_equal(perm.astype(np.int64), exp)
    # the radix restatement with its 11-bit (or 8-bit) digits gives the same permutation
    _, rperm = oracle.radixsort(bits, flags, want_payload=True)
    assert np.array_equal(rperm, perm)


# ---- the radix algorithm against the merge sort and numpy ----------------------------------

def _random_keys(w, n, rng, kind):
    hi = np.iinfo(UINT[w]).max
    if kind == "uniform":
        return rng.integers(0, hi, n, dtype=UINT[w], endpoint=True)
    if kind == "dups":
        pool = rng.integers(0, hi, 17, dtype=UINT[w], endpoint=True)
        return pool[rng.integers(0, 17, n)]
    if kind == "lowbits":
        return (rng.integers(0, 1 << 12, n, dtype=np.uint64) | np.uint64(0xA5A5A5A5A5A5A5A5 & hi & ~0xFFF)).astype(UINT[w])
    if kind == "sorted":
        return np.sort(rng.integers(0, hi, n, dtype=UINT[w], endpoint=True))
    if kind == "equal":
        return np.full(n, rng.integers(0, hi, dtype=UINT[w], endpoint=True), UINT[w])
    raise AssertionError(kind)


@pytest.mark.parametrize("w", [8, 16, 32, 64])
@pytest.mark.parametrize("kind", ["uniform", "dups", "lowbits", "sorted", "equal"])
def test_radix_equals_mergesort_all_modes(w, kind):
    rng = np.random.default_rng(w + len(kind))
    for n in (0, 1, 2, 3, 257, 3001):
        keys = _random_keys(w, n, rng, kind)
        for mode in MODES:
            for d in DIRECTIONS:
                flags = oracle.decode(oracle.I8, mode, d, True)
                a, pa = oracle.radixsort(keys, flags, want_payload=True)
                b, pb = oracle.mergesort(keys, flags)
                assert a.tolist() == b.tolist(), (n, mode, d)
                assert pa.tolist() == pb.tolist(), (n, mode, d)
                # keys-only path (8-bit count fill included)
                c, _ = oracle.radixsort(keys, flags)
                assert c.tolist() == a.tolist()


@pytest.mark.parametrize("w", [16, 32, 64])
def test_radix_matches_numpy_stable_argsort(w):
    rng = np.random.default_rng(100 + w)
    keys = _random_keys(w, 50000, rng, "dups")
    for mode in MODES:
        flags = oracle.decode(oracle.I8, mode, 0)
        k = np.array([oracle.sortkey(w, flags, int(x)) for x in keys[:2000]], dtype=np.uint64)
        exp = np.argsort(k, kind="stable")
        _, perm = oracle.radixsort(keys[:2000], flags, want_payload=True)
        assert np.array_equal(perm.astype(np.int64), exp)


def test_invariants_and_payload_passthrough():
    rng = np.random.default_rng(5)
    keys = _random_keys(64, 20000, rng, "dups")
    payload = rng.integers(0, 1 << 32, keys.size, dtype=np.uint64).astype(np.uint32)
    for d in DIRECTIONS:
        flags = oracle.decode(oracle.F64, 64, d, True)
        out, perm = oracle.radixsort(keys, flags, want_payload=True)
        out2, pay = oracle.radixsort(keys, flags, payload=payload)
        assert np.array_equal(np.sort(out), np.sort(keys))           # permutation
        assert np.array_equal(keys[perm], out)
        assert np.array_equal(payload[perm], pay)                    # payload travels
        k = np.array([oracle.sortkey(64, flags, int(x)) for x in out], dtype=np.uint64)
        step = np.diff(k.astype(np.float64))
        assert (step <= 0).all() if d & 1 else (step >= 0).all()     # monotone keys
        same = k[1:] == k[:-1]
        order = np.diff(perm.astype(np.int64))[same]
        assert (order < 0).all() if d & 2 else (order > 0).all()     # stability / reversal


def test_pass_skip_counts():
    """One executed pass per non-uniform reference digit (11/11/11/11/11/9 bits for 64-bit
    keys, H:1848-1869); all-equal keys execute none."""
    keys = np.full(1000, 0x1234, np.uint64)
    assert oracle.passes_executed(keys, 0) == 0
    keys[::2] = 0x1234 + (1 << 11)
    assert oracle.passes_executed(keys, 0) == 1
    keys[::3] |= np.uint64(1 << 60)
    assert oracle.passes_executed(keys, 0) == 2


def test_histogram8_matches_bincount():
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 1 << 32, 10000, dtype=np.uint64).astype(np.uint32)
    for mode in MODES:
        for d in (0, 1):
            flags = oracle.decode(oracle.I8, mode, d)
            h = oracle.histogram8(keys, flags)
            k = np.array([oracle.sortkey(32, flags, int(x)) for x in keys], dtype=np.uint64)
            if d:
                k ^= np.uint64(0xFFFFFFFF)
            for p in range(4):
                exp = np.bincount(((k >> np.uint64(8 * p)) & np.uint64(255)).astype(np.int64), minlength=256)
                assert np.array_equal(h[p].astype(np.int64), exp)


@pytest.mark.parametrize("w", [16, 32, 64])
def test_multithreaded_matches(w):
    rng = np.random.default_rng(w)
    keys = _random_keys(w, 300000, rng, "uniform")
    for mode, d in ((0, 0), (6, 3), (3, 1), (5, 2)):
        flags = oracle.decode(oracle.I8, mode, d)
        a, _ = oracle.radixsort(keys, flags)
        b = oracle.radixsort_mt(keys, flags, 4)
        assert np.array_equal(a, b)


def test_indirect_and_segmented():
    rng = np.random.default_rng(11)
    n, stride, off = 5000, 32, 8
    rec = rng.integers(0, 256, n * stride, dtype=np.uint8)
    rec.view(np.uint64)[1::4] &= np.uint64(0xFF)  # many duplicate keys at offset 8
    idx = rng.permutation(n).astype(np.uint32)
    flags = oracle.decode(oracle.I64, 64, 3, True)
    got = oracle.sort_indirect(rec, stride, off, idx, np.int64, flags)
    keys = rec.view(np.uint64)[1::4][idx]
    _, perm = oracle.mergesort(keys, flags)
    assert np.array_equal(got, idx[perm])

    keys = rng.integers(-2**31, 2**31, 20000, dtype=np.int64).astype(np.int32)
    seg = np.array([0, 0, 1, 17, 4000, 4001, 20000], np.uint64)
    flags = oracle.decode(oracle.I32, 64, 0)
    out = oracle.sort_segmented(keys, seg, flags)
    for b, e in zip(seg[:-1], seg[1:]):
        assert np.array_equal(out[int(b):int(e)], np.sort(keys[int(b):int(e)], kind="stable"))


def test_vector_sortkeys_and_order_checker():
    """sortkeys equals the scalar sortkey; is_stable_sort accepts the oracle's own result and
    rejects a broken permutation, a swapped tie and a swapped pair of distinct keys."""
    rng = np.random.default_rng(77)
    for w, dt in ((8, np.uint8), (16, np.uint16), (32, np.uint32), (64, np.uint64)):
        keys = rng.integers(0, np.iinfo(dt).max, 500, dtype=dt, endpoint=True)
        for flags in range(32):
            v = oracle.sortkeys(keys, flags)
            assert all(int(v[i]) == oracle.sortkey(w, flags, int(keys[i])) for i in range(0, 500, 7))
    keys = rng.integers(0, 50, 4000).astype(np.uint32)
    for flags in (0, 8, 16, 24, 2, 26):
        out, perm = oracle.radixsort(keys, flags, want_payload=True)
        assert oracle.is_stable_sort(keys, out, perm, flags)
        bad = perm.copy()
        bad[0] = bad[1]
        assert not oracle.is_stable_sort(keys, keys[bad], bad, flags)
        sk = oracle.sortkeys(out, flags)
        i = int(np.nonzero(sk[:-1] == sk[1:])[0][0])          # swap two equal keys: the tie order breaks
        p2 = perm.copy()
        p2[i], p2[i + 1] = p2[i + 1], p2[i]
        assert not oracle.is_stable_sort(keys, keys[p2], p2, flags)
        j = int(np.nonzero(sk[:-1] != sk[1:])[0][0])          # swap two different keys: the order breaks
        p3 = perm.copy()
        p3[j], p3[j + 1] = p3[j + 1], p3[j]
        assert not oracle.is_stable_sort(keys, keys[p3], p3, flags)
