"""Pins of the oracle's keys wider than 64 bits (rsbd8o_sortkey_wide, rsbd8o_radixsort_wide), on the CPU.

- the reference's own known answers for the 80-bit layouts (multiplatform/main.cpp:21-178, kept as
  data in tests/golden/kat_wide.json), whose padding bytes hold arbitrary values;
- the floating-point modes of the 16-byte layout against numpy's longdouble (the x86-64 80-bit
  extended type): value order, |value| order, and the documented tiered and inside-out sequences
  (H:241-247) restated as Python sort keys over the values' signs and magnitudes;
- the integer modes of all layouts against Python integers: unsigned and two's-complement order,
  |value| order, and the tiered sequence 0, -1, 1, -2, 2, ... (H:247-248);
- descending order and the tie reversal as the reversed stable orders;
- the padding of the 12- and 16-byte layouts never changing the order.
"""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat_wide.json")))
LAYOUT = {"ld80": oracle.LD80, "ld96": oracle.LD96, "ld128": oracle.LD128}
ABS, SIGN, FLTP, DESC, REV = oracle.ABS, oracle.SIGN, oracle.FLTP, oracle.DESC, oracle.REV


def records(keytype, lo, hi, rng=None):
    """Records of a wide layout from 64-bit low words and high words (the 16-bit sign/exponent for
    the 80-bit layouts; random padding above it when rng is given)."""
    rb, vb = oracle.WIDE_LAYOUT[keytype]
    n = len(lo)
    out = np.zeros((n, rb), np.uint8)
    out[:, :8] = np.asarray(lo, np.uint64).view(np.uint8).reshape(n, 8)
    if vb == 80:
        out[:, 8:10] = np.asarray(hi, np.uint64).astype(np.uint16).view(np.uint8).reshape(n, 2)
        if rng is not None and rb > 10:
            out[:, 10:] = rng.integers(0, 256, (n, rb - 10), dtype=np.uint8)
    else:
        out[:, 8:16] = np.asarray(hi, np.uint64).view(np.uint8).reshape(n, 8)
    return out


def stable_order(keys, desc=False, rev=False):
    """The stable order of Python sort keys: ascending (descending), ties in source order or reversed."""
    idx = list(range(len(keys)))
    if rev:
        idx.reverse()
    return np.array(sorted(idx, key=lambda i: keys[i], reverse=False) if not desc else
                    _desc_stable(idx, keys), np.int64)


def _desc_stable(idx, keys):
    # descending with ties kept in the given order: group by key from the top
    return sorted(idx, key=lambda i: _Neg(keys[i]))


class _Neg:
    __slots__ = ("v",)

    def __init__(self, v):
        self.v = v

    def __lt__(self, o):
        return o.v < self.v


@pytest.mark.parametrize("case", KAT["cases"], ids=lambda c: c["name"])
def test_wide_known_answers(case):
    kt = LAYOUT[case["layout"]]
    lo = [int(x) for x in case["input_mantissa"]]
    se = [int(x) for x in case["input_signexponent"]]
    rb = oracle.WIDE_LAYOUT[kt][0]
    recs = np.zeros((7, rb), np.uint8)
    for i in range(7):
        recs[i, :8] = np.frombuffer(int(lo[i]).to_bytes(8, "little"), np.uint8)
        recs[i, 8:] = np.frombuffer(int(se[i]).to_bytes(rb - 8, "little"), np.uint8)
    flags = oracle.decode(kt, case["mode"], case["direction"], case["indirect"])
    assert flags & (SIGN | FLTP) == SIGN | FLTP and not flags & ABS
    perm = oracle.radixsort_wide(kt, flags, recs)
    assert perm.tolist() == case["expected_perm"]


def _finite_longdoubles(rng, n):
    """Canonical 80-bit patterns: zeros and subnormals (exponent 0, integer bit clear), normals
    (integer bit set), infinities; with repeated values and both zeros."""
    exp = rng.integers(0, 0x7FFF, n).astype(np.uint64)
    exp[rng.random(n) < 0.05] = 0
    exp[rng.random(n) < 0.02] = 0x7FFF
    mant = rng.integers(0, 1 << 63, n, dtype=np.uint64)
    mant = np.where(exp > 0, mant | np.uint64(1 << 63), mant & np.uint64((1 << 63) - 1))
    mant[exp == 0x7FFF] = np.uint64(1 << 63)  # infinity
    zero = rng.random(n) < 0.03
    mant[zero], exp[zero] = 0, 0
    sign = (rng.random(n) < 0.5).astype(np.uint64) << np.uint64(15)
    se = exp | sign
    dup = rng.integers(0, n, n // 10)
    mant[dup[1:]], se[dup[1:]] = mant[dup[0]], se[dup[0]]
    return mant, se


@pytest.mark.skipif(np.dtype(np.longdouble).itemsize != 16 or np.finfo(np.longdouble).nmant != 63,
                    reason="numpy's longdouble is not the x86 80-bit extended type here")
@pytest.mark.parametrize("direction", [0, 1, 2, 3])
def test_wide_float_modes_against_numpy_longdouble(direction):
    rng = np.random.default_rng(80 + direction)
    n = 3000
    mant, se = _finite_longdoubles(rng, n)
    recs = records(oracle.LD128, mant, se)
    vals = recs.view(np.longdouble).ravel()
    assert np.isfinite(vals[(se & 0x7FFF) != 0x7FFF]).all()
    # ranks of the values and of their magnitudes, equal values sharing a rank (-0 == +0)
    order = np.argsort(vals, kind="stable")
    same = np.concatenate([[False], vals[order][1:] == vals[order][:-1]])
    rk = np.cumsum(~same) - 1
    vrank = np.empty(n, np.int64)
    vrank[order] = rk
    neg = np.signbit(vals)
    mag_order = np.argsort(np.abs(vals), kind="stable")
    msame = np.concatenate([[False], np.abs(vals)[mag_order][1:] == np.abs(vals)[mag_order][:-1]])
    mrank = np.empty(n, np.int64)
    mrank[mag_order] = np.cumsum(~msame) - 1
    desc, rev = bool(direction & 1), bool(direction & 2)
    cases = {
        SIGN | FLTP: [int(r) for r in vrank],                                     # value order, -0 == +0
        ABS | SIGN | FLTP: [int(r) for r in mrank],                               # |value| order
        ABS | FLTP: [(int(mrank[i]), 0 if neg[i] else 1) for i in range(n)],      # -0, +0, -min, +min, ...
        FLTP: [(1, -int(mrank[i])) if neg[i] else (0, int(mrank[i])) for i in range(n)],  # +0 .. +inf, -inf .. -0
    }
    for fl, keys in cases.items():
        flags = fl | (DESC if desc else 0) | (REV if rev else 0)
        got = oracle.radixsort_wide(oracle.LD128, flags, recs)
        assert np.array_equal(got, stable_order(keys, desc, rev)), fl


@pytest.mark.parametrize("keytype", [oracle.LD80, oracle.LD96, oracle.LD128, oracle.U128, oracle.I128])
@pytest.mark.parametrize("direction", [0, 1, 2, 3])
def test_wide_integer_modes_against_python_ints(keytype, direction):
    rng = np.random.default_rng(keytype * 10 + direction)
    n = 2000
    w = oracle.WIDE_LAYOUT[keytype][1]
    lo = rng.integers(0, 1 << 64, n, dtype=np.uint64, endpoint=False)
    hi = rng.integers(0, 1 << 64, n, dtype=np.uint64, endpoint=False)
    if w == 80:
        hi &= np.uint64(0xFFFF)
    # specials: zero, the minimum, -1, small magnitudes of both signs, repeats
    sp = rng.random(n) < 0.1
    lo[sp], hi[sp] = 0, 0
    top = np.uint64((1 << (w - 64 - 1)))
    mn = rng.random(n) < 0.03
    lo[mn], hi[mn] = 0, top
    m1 = rng.random(n) < 0.03
    lo[m1], hi[m1] = np.uint64((1 << 64) - 1), np.uint64((1 << (w - 64)) - 1)
    recs = records(keytype, lo, hi, rng)
    u = [(int(hi[i]) << 64) | int(lo[i]) for i in range(n)]
    s = [x - (1 << w) if x >> (w - 1) else x for x in u]
    desc, rev = bool(direction & 1), bool(direction & 2)
    cases = {
        0: u,                                                          # unsigned
        SIGN: s,                                                       # two's complement
        ABS | SIGN: [abs(x) for x in s],                               # |x|, ties x and -x
        ABS: [2 * abs(x) - 1 if x < 0 else 2 * x for x in s],          # 0, -1, 1, -2, 2, ...
    }
    for fl, keys in cases.items():
        flags = fl | (DESC if desc else 0) | (REV if rev else 0)
        got = oracle.radixsort_wide(keytype, flags, recs)
        assert np.array_equal(got, stable_order(keys, desc, rev)), (keytype, fl)


def test_wide_sortkey_words_and_padding():
    """The key words for the 80-bit layouts use 80 bits only; the padding never reaches the key."""
    rng = np.random.default_rng(5)
    mant, se = _finite_longdoubles(rng, 500)
    for kt in (oracle.LD96, oracle.LD128):
        a = records(kt, mant, se, rng)
        b = records(kt, mant, se, np.random.default_rng(6))
        assert not np.array_equal(a, b)
        for fl in range(8):
            ka, kb = oracle.sortkey_wide(kt, fl, a), oracle.sortkey_wide(kt, fl, b)
            assert ka == kb
            assert all(k < (1 << 80) for k in ka)
            assert np.array_equal(oracle.radixsort_wide(kt, fl, a), oracle.radixsort_wide(kt, fl, b))
        assert oracle.sortkey_wide(kt, 0, a) == oracle.sortkey_wide(oracle.LD80, 0, records(oracle.LD80, mant, se))
