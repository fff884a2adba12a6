"""CPU oracle for the rsbd8 radix sort -- TEST INFRASTRUCTURE ONLY.

ctypes front end over ``oracle/_build/librsbd8_oracle.so`` (built from ``rsbd8_oracle.c`` by
``oracle/Makefile``).  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this package, and only as the checker / the timed CPU baseline;
the product package ``radixsortbidi8_amd`` never imports it.

Every function restates part of the reference (Radixsortbidi8.hpp = H):
  decode        H:59778-59904 (mode/direction decoding)
  sortkey       H:10000-10027, H:3375-3412, H:11642-11801 (key filters and virtual bit flips)
  radixsort     H:32107-32473 (counting pass), H:11626-11809 (offsets, pass skip),
                H:31649-31982 (bidirectional scatter), H:48769-48957 (8-bit count fill)
  mergesort     an independent stable merge sort on (sortkey, logical index), H:252-275
How each is pinned is listed at the top of rsbd8_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ABS, SIGN, FLTP, DESC, REV = 1, 2, 4, 8, 16

# key type codes (shared with include/rsbd8_hip.h)
U8, U16, U32, U64, I8, I16, I32, I64, F32, F64 = range(10)
_DTYPE_CODE = {
    np.dtype(np.uint8): U8, np.dtype(np.uint16): U16, np.dtype(np.uint32): U32,
    np.dtype(np.uint64): U64, np.dtype(np.int8): I8, np.dtype(np.int16): I16,
    np.dtype(np.int32): I32, np.dtype(np.int64): I64, np.dtype(np.float32): F32,
    np.dtype(np.float64): F64,
}
_UNSIGNED = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "librsbd8_oracle.so")
_lib = None


def build() -> str:
    """Compile the oracle library with oracle/Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_size, c_vp, c_u = ctypes.c_size_t, ctypes.c_void_p, ctypes.c_uint
        L.rsbd8o_decode.argtypes = [ctypes.c_int, c_u, c_u, ctypes.c_int]
        L.rsbd8o_decode.restype = c_u
        L.rsbd8o_sortkey.argtypes = [c_u, c_u, ctypes.c_uint64]
        L.rsbd8o_sortkey.restype = ctypes.c_uint64
        L.rsbd8o_radixsort.argtypes = [c_u, c_u, c_size, c_vp, c_vp, c_vp, c_vp]
        L.rsbd8o_radixsort.restype = ctypes.c_int
        L.rsbd8o_mergesort.argtypes = [c_u, c_u, c_size, c_vp, c_vp, c_vp]
        L.rsbd8o_mergesort.restype = ctypes.c_int
        L.rsbd8o_sort_indirect.argtypes = [c_u, c_u, c_size, c_vp, c_size, c_size, c_vp, c_vp]
        L.rsbd8o_sort_indirect.restype = ctypes.c_int
        L.rsbd8o_sort_segmented.argtypes = [c_u, c_u, c_size, c_vp, c_vp, c_vp]
        L.rsbd8o_sort_segmented.restype = ctypes.c_int
        L.rsbd8o_histogram8.argtypes = [c_u, c_u, c_size, c_vp, c_vp]
        L.rsbd8o_histogram8.restype = None
        L.rsbd8o_sortkeys.argtypes = [c_u, c_u, c_size, c_vp, c_vp]
        L.rsbd8o_sortkeys.restype = None
        L.rsbd8o_radixsort_mt.argtypes = [c_u, c_u, c_size, c_vp, c_vp, c_u]
        L.rsbd8o_radixsort_mt.restype = ctypes.c_int
        L.rsbd8o_sortkey_wide.argtypes = [ctypes.c_int, c_u, c_vp, c_vp, c_vp]
        L.rsbd8o_sortkey_wide.restype = None
        L.rsbd8o_radixsort_wide.argtypes = [ctypes.c_int, c_u, c_size, c_vp, c_size, c_vp]
        L.rsbd8o_radixsort_wide.restype = ctypes.c_int
        _lib = L
    return _lib


# keys wider than 64 bits (H:1061-1804, H:2186-2404): the 80-bit extended float in 10-, 12- and
# 16-byte records, and 128-bit integers; bytes per record and value bits
LD80, LD96, LD128, U128, I128 = 10, 11, 12, 13, 14
WIDE_LAYOUT = {LD80: (10, 80), LD96: (12, 80), LD128: (16, 80), U128: (16, 128), I128: (16, 128)}


def sortkey_wide(keytype: int, flags: int, recs: np.ndarray):
    """The 128-bit sort keys (as Python ints) of the records in recs (uint8, n * record bytes)."""
    rb = WIDE_LAYOUT[keytype][0]
    b = np.ascontiguousarray(recs, dtype=np.uint8).reshape(-1, rb)
    lo, hi = ctypes.c_uint64(), ctypes.c_uint64()
    out = []
    for row in b:
        lib().rsbd8o_sortkey_wide(keytype, flags, row.ctypes.data_as(ctypes.c_void_p), ctypes.byref(lo), ctypes.byref(hi))
        out.append((hi.value << 64) | lo.value)
    return out


def radixsort_wide(keytype: int, flags: int, recs: np.ndarray) -> np.ndarray:
    """The stable permutation of the records in recs (uint8, n * record bytes) by their sort keys."""
    rb = WIDE_LAYOUT[keytype][0]
    b = np.ascontiguousarray(recs, dtype=np.uint8).ravel()
    n = b.size // rb
    perm = np.empty(n, np.uint32)
    assert lib().rsbd8o_radixsort_wide(keytype, flags, n, b.ctypes.data_as(ctypes.c_void_p), rb,
                                       perm.ctypes.data_as(ctypes.c_void_p)) == 0
    return perm


def keytype_of(dtype) -> int:
    return _DTYPE_CODE[np.dtype(dtype)]


def decode(keytype: int, mode: int, direction: int, indirect: bool = False) -> int:
    return int(lib().rsbd8o_decode(keytype, mode, direction, int(bool(indirect))))


def sortkey(width_bits: int, flags: int, x: int) -> int:
    return int(lib().rsbd8o_sortkey(width_bits, flags, x))


def sortkeys(keys: np.ndarray, flags: int) -> np.ndarray:
    """rsbd8o_sortkey of every key (the descending bit not applied), as uint64."""
    kb, w = _bits(keys)
    out = np.empty(kb.size, np.uint64)
    lib().rsbd8o_sortkeys(w, flags, kb.size, _ptr(kb), _ptr(out))
    return out


def is_stable_sort(keys_in: np.ndarray, keys_out: np.ndarray, perm: np.ndarray, flags: int) -> bool:
    """Whether keys_out with source permutation perm is exactly the reference's stable sort of
    keys_in under flags, checked through properties that hold at any size: perm is a permutation,
    keys_out is keys_in gathered by perm (bit for bit), and consecutive outputs are ordered by
    (sort key ascending, or descending with the descending bit; then source index ascending, or
    descending with the tie reversal), the order rsbd8o_mergesort sorts by."""
    n = keys_in.size
    perm = np.ascontiguousarray(perm).astype(np.int64, copy=False)
    if keys_out.size != n or perm.size != n:
        return False
    if n == 0:
        return True
    if perm.min() < 0 or perm.max() >= n:
        return False
    seen = np.zeros(n, bool)
    seen[perm] = True
    if not seen.all():
        return False
    kin, _ = _bits(keys_in)
    kout, _ = _bits(keys_out)
    if not np.array_equal(kin[perm], kout):
        return False
    sk = sortkeys(keys_out, flags)
    a, b = sk[:-1], sk[1:]
    pa, pb = perm[:-1], perm[1:]
    strict = (a > b) if (flags & 8) else (a < b)
    ties = (pa > pb) if (flags & 16) else (pa < pb)
    return bool(np.all(strict | ((a == b) & ties)))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _bits(keys: np.ndarray):
    keys = np.ascontiguousarray(keys)
    return keys.view(_UNSIGNED[keys.dtype.itemsize]), keys.dtype.itemsize * 8


def radixsort(keys: np.ndarray, flags: int, payload: np.ndarray | None = None,
              want_payload: bool = False):
    """Stable LSD radix sort (the reference's algorithm). Returns (sorted keys, payload or None).
    With want_payload and no payload the payload is the source index."""
    kb, w = _bits(keys)
    out = np.empty_like(kb)
    pin = None if payload is None else np.ascontiguousarray(payload, dtype=np.uint32)
    pout = np.empty(kb.size, np.uint32) if (want_payload or pin is not None) else None
    r = lib().rsbd8o_radixsort(w, flags, kb.size, _ptr(kb), _ptr(out), _ptr(pin), _ptr(pout))
    if r < 0:
        raise MemoryError("oracle radixsort failed")
    return out.view(keys.dtype), pout


def passes_executed(keys: np.ndarray, flags: int) -> int:
    kb, w = _bits(keys)
    out = np.empty_like(kb)
    return int(lib().rsbd8o_radixsort(w, flags, kb.size, _ptr(kb), _ptr(out), None, None))


def mergesort(keys: np.ndarray, flags: int):
    """Independent stable merge sort. Returns (sorted keys, source permutation)."""
    kb, w = _bits(keys)
    out = np.empty_like(kb)
    perm = np.empty(kb.size, np.uint32)
    if lib().rsbd8o_mergesort(w, flags, kb.size, _ptr(kb), _ptr(out), _ptr(perm)) < 0:
        raise MemoryError("oracle mergesort failed")
    return out.view(keys.dtype), perm


def sort_indirect(records: np.ndarray, stride: int, offset: int, idx: np.ndarray,
                  key_dtype, flags: int) -> np.ndarray:
    """Sort u32 record indices by the key of type key_dtype at byte offset in each record."""
    rec = np.ascontiguousarray(records).view(np.uint8)
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    out = np.empty_like(idx)
    w = np.dtype(key_dtype).itemsize * 8
    if lib().rsbd8o_sort_indirect(w, flags, idx.size, _ptr(rec), stride, offset, _ptr(idx), _ptr(out)) < 0:
        raise MemoryError("oracle sort_indirect failed")
    return out


def sort_segmented(keys: np.ndarray, seg: np.ndarray, flags: int) -> np.ndarray:
    kb, w = _bits(keys)
    seg = np.ascontiguousarray(seg, dtype=np.uint64)
    out = np.empty_like(kb)
    if lib().rsbd8o_sort_segmented(w, flags, seg.size - 1, _ptr(seg), _ptr(kb), _ptr(out)) < 0:
        raise ValueError("oracle sort_segmented failed")
    return out.view(keys.dtype)


def histogram8(keys: np.ndarray, flags: int) -> np.ndarray:
    kb, w = _bits(keys)
    counts = np.zeros((w // 8, 256), np.uint64)
    lib().rsbd8o_histogram8(w, flags, kb.size, _ptr(kb), _ptr(counts))
    return counts


def radixsort_mt(keys: np.ndarray, flags: int, threads: int, out: np.ndarray | None = None):
    kb, w = _bits(keys)
    o = np.empty_like(kb) if out is None else out.view(kb.dtype)
    if lib().rsbd8o_radixsort_mt(w, flags, kb.size, _ptr(kb), _ptr(o), threads) < 0:
        raise MemoryError("oracle radixsort_mt failed")
    return o.view(keys.dtype)
