"""GPU parity of the HIP sort against the CPU oracle, through the C ABI.

Bit-exact comparisons (integer/byte work): sorted keys, and the source permutation wherever a
payload or pointers travel. Covers every key width, all eight modes and four directions, the
reference's known answers, degenerate sizes, uniform digits (pass skipping with odd and even
active-pass counts, in place and copying), payloads, and the index/pointer indirection forms.
"""
import json
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
r = pytest.importorskip("radixsortbidi8_amd")

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "kat_reference.json")))
UINT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
SIGNED_DT = {1: np.int8, 2: np.int16, 4: np.int32, 8: np.int64}
FLOAT_DT = {4: np.float32, 8: np.float64}


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def gen(nbytes, n, kind, rng):
    hi = np.iinfo(UINT[nbytes]).max
    if kind == "uniform":
        return rng.integers(0, hi, n, dtype=UINT[nbytes], endpoint=True)
    if kind == "dups":
        pool = rng.integers(0, hi, 23, dtype=UINT[nbytes], endpoint=True)
        return pool[rng.integers(0, 23, n)]
    if kind == "lowbyte":     # one varying byte: one active pass
        return (np.uint64(0x9E3779B97F4A7C15 & hi & ~0xFF) | rng.integers(0, 256, n, dtype=np.uint64)).astype(UINT[nbytes])
    if kind == "twobytes":    # two varying bytes (bytes 0 and 2 when wide enough): two active passes
        v = rng.integers(0, 256, n, dtype=np.uint64) | (rng.integers(0, 256, n, dtype=np.uint64) << np.uint64(16 if nbytes > 2 else 8))
        return (np.uint64(0x5555555555555555 & hi & ~0xFF00FF) | v).astype(UINT[nbytes])
    if kind == "equal":
        return np.full(n, rng.integers(0, hi, dtype=UINT[nbytes], endpoint=True), UINT[nbytes])
    if kind == "sorted":
        return np.sort(rng.integers(0, hi, n, dtype=UINT[nbytes], endpoint=True))
    if kind == "specials":    # sign/zero/inf/NaN patterns mixed with random ones
        msb = 1 << (8 * nbytes - 1)
        pool = [0, msb, 1, msb | 1, msb - 1, hi]
        if nbytes == 4:
            pool += [0x7F800000, 0xFF800000, 0x7FC00000, 0xFFC00000, 0x7F800001, 0x00000001, 0x80000001]
        if nbytes == 8:
            pool += [0x7FF0000000000000, 0xFFF0000000000000, 0x7FF8000000000000, 0xFFF8000000000000]
        a = rng.integers(0, hi, n, dtype=UINT[nbytes], endpoint=True)
        m = rng.random(n) < 0.3
        a[m] = np.array(pool, dtype=UINT[nbytes])[rng.integers(0, len(pool), int(m.sum()))]
        return a
    raise AssertionError(kind)


def expected(keys, flags):
    out, perm = oracle.radixsort(keys, flags, want_payload=True)
    return out, perm


def typed(bits, kt):
    nbytes = bits.dtype.itemsize
    if kt in (r.F32, r.F64):
        return bits.view(FLOAT_DT[nbytes])
    if kt >= r.I8:
        return bits.view(SIGNED_DT[nbytes])
    return bits


KEYTYPES = {1: [r.U8, r.I8], 2: [r.U16, r.I16], 4: [r.U32, r.I32, r.F32], 8: [r.U64, r.I64, r.F64]}
TORCH_DT = {r.U8: torch.uint8, r.I8: torch.int8, r.U16: torch.uint16, r.I16: torch.int16, r.U32: torch.uint32,
            r.I32: torch.int32, r.F32: torch.float32, r.U64: torch.uint64, r.I64: torch.int64,
            r.F64: torch.float64}


def combos(nbytes):
    """(keytype, mode) pairs: the native modes per key type, the explicit modes once (they do not
    depend on the type)."""
    out = [(kt, m) for kt in KEYTYPES[nbytes] for m in (64, 65, 66)]
    return out + [(KEYTYPES[nbytes][0], m) for m in range(8)]


@pytest.fixture(params=[1, 2, 4], ids=["onesweep", "countscan", "msd2"])
def algo(request):
    """Runs a test under each pass algorithm (onesweep look-back, count + chained scan + scatter,
    two MSD partitions + per-bucket LDS sort, which keys-only 4/8-byte sorts take and the others
    replace by count + scan)."""
    prev = r.set_algorithm(request.param)
    yield request.param
    r.set_algorithm(prev)


# ---- known answers ---------------------------------------------------------------------------

@pytest.mark.parametrize("case", KAT["cases"], ids=lambda c: c["name"])
def test_known_answers_on_gpu(case):
    dt = np.dtype(case["dtype"])
    nb = dt.itemsize
    keys = (np.array(case["input_bits"], dtype=UINT[nb]).view(dt) if "input_bits" in case
            else np.array(case["input"], dtype=dt))
    kt = r.keytype_of(dt)
    if not case["indirect"]:
        d_in, d_out = dev(keys), dev(np.zeros_like(keys))
        if case["api"] == "radixsort":
            assert r.radixsort(d_in, case["direction"], case["mode"])
            got = host(d_in)
        else:
            assert r.radixsortcopy(d_in, d_out, case["direction"], case["mode"])
            got = host(d_out)
        exp = (np.array(case["expected_bits"], dtype=UINT[nb]) if "expected_bits" in case
               else np.array(case["expected"], dtype=dt).view(UINT[nb]))
        assert got.view(UINT[nb]).tolist() == exp.tolist()
        return
    # indirection: device pointers to the keys, sorted by the key they point at
    d_keys = dev(keys)
    ptrs = torch.arange(keys.size, dtype=torch.int64, device="cuda") * nb + d_keys.data_ptr()
    out = torch.zeros_like(ptrs)
    ws = r.Workspace(kt, keys.size, r.LAYOUT_INDIRECT)
    r.radixsortcopynoalloc_pointers(ptrs, out, ws, dt, 0, case["direction"], case["mode"])
    perm = ((host(out) - d_keys.data_ptr()) // nb).tolist()
    if "expected_perm" in case:
        assert perm == case["expected_perm"]
    if "expected_bits" in case:
        assert keys.view(UINT[nb])[perm].tolist() == case["expected_bits"]


# ---- parity with the oracle --------------------------------------------------------------------

@pytest.mark.parametrize("nbytes", [1, 2, 4, 8])
@pytest.mark.parametrize("kind", ["uniform", "dups", "specials", "lowbyte", "twobytes", "equal", "sorted"])
def test_copy_all_modes_directions(nbytes, kind, algo):
    rng = np.random.default_rng(nbytes * 131 + len(kind))
    for n in (2, 3, 1000, 70001):
        bits = gen(nbytes, n, kind, rng)
        ws = r.Workspace(r.U8, n)
        d_in = dev(bits)
        d_out, d_buf = torch.empty_like(d_in), torch.empty_like(d_in)
        cache = {}
        for kt, mode in combos(nbytes):
            for d in range(4):
                flags = r.decode(kt, mode, d)
                t = TORCH_DT[kt]
                r.radixsortcopynoalloc(d_in.view(t), d_out.view(t), d_buf.view(t), ws, d, mode)
                if flags not in cache:
                    cache[flags] = expected(bits, flags)[0]
                got = host(d_out)
                assert np.array_equal(got, cache[flags]), (n, kt, mode, d)
        assert np.array_equal(host(d_in), bits)  # input untouched


@pytest.mark.parametrize("nbytes", [1, 2, 4, 8])
@pytest.mark.parametrize("kind", ["uniform", "lowbyte", "twobytes", "equal"])
def test_inplace_movetobuffer(nbytes, kind, algo):
    """radixsortnoalloc in both output placements; odd and even active-pass counts exercise the
    staging copy the in-place plan needs."""
    rng = np.random.default_rng(7 * nbytes + len(kind))
    n = 50003
    bits = gen(nbytes, n, kind, rng)
    ws = r.Workspace(r.U8, n)
    kt = KEYTYPES[nbytes][-1]
    for mode in (64, 65, 66, 0, 5):
        for d in range(4):
            flags = r.decode(kt, mode, d)
            exp, _ = expected(bits, flags)
            for move in (False, True):
                a = dev(bits).view(TORCH_DT[kt])
                b = torch.empty_like(a)
                r.radixsortnoalloc(a, b, ws, move, d, mode)
                got = host(b if move else a).view(bits.dtype)
                assert np.array_equal(got, exp), (mode, d, move)


def test_allocating_forms_host_and_device():
    rng = np.random.default_rng(3)
    for nbytes, kt in ((4, r.F32), (8, r.I64), (2, r.U16), (1, r.I8)):
        bits = gen(nbytes, 12345, "specials", rng)
        keys = typed(bits, kt)
        for mode in (64, 65):
            for d in range(4):
                exp, _ = expected(bits, r.decode(kt, mode, d))
                out = np.empty_like(keys)
                assert r.radixsortcopy(keys, out, d, mode)          # host arrays, staged
                assert np.array_equal(out.view(bits.dtype), exp)
                h = keys.copy()
                assert r.radixsort(h, d, mode)                       # host, in place
                assert np.array_equal(h.view(bits.dtype), exp)
                t = dev(keys)
                assert r.radixsort(t, d, mode)                       # device, in place
                assert np.array_equal(host(t).view(bits.dtype), exp)


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4])
def test_every_tile_config(cfg, algo):
    rng = np.random.default_rng(cfg)
    prev = r.set_tile_config(cfg)
    try:
        for nbytes in (2, 4, 8):
            n = 300007
            bits = gen(nbytes, n, "dups" if cfg % 2 else "uniform", rng)
            kt = KEYTYPES[nbytes][-1]
            ws = r.Workspace(kt, n)
            a = dev(bits)
            o, b = torch.empty_like(a), torch.empty_like(a)
            t = TORCH_DT[kt]
            for d in (0, 3):
                r.radixsortcopynoalloc(a.view(t), o.view(t), b.view(t), ws, d, 64)
                exp, _ = expected(bits, r.decode(kt, 64, d))
                assert np.array_equal(host(o), exp)
    finally:
        r.set_tile_config(prev)


@pytest.mark.parametrize("nbytes", [1, 2, 4, 8])
def test_pairs_payload(nbytes, algo):
    rng = np.random.default_rng(40 + nbytes)
    n = 99991
    for kind in ("dups", "specials", "equal"):
        bits = gen(nbytes, n, kind, rng)
        payload = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        kt = KEYTYPES[nbytes][-1]
        ws = r.Workspace(kt, n, r.LAYOUT_PAIRS)
        k = dev(typed(bits, kt))
        ko, kb = torch.empty_like(k), torch.empty_like(k)
        p = dev(payload)
        po, pb = torch.empty_like(p), torch.empty_like(p)
        for mode in (64, 65, 66):
            for d in range(4):
                flags = r.decode(kt, mode, d, True)
                exp_k, exp_p = oracle.radixsort(bits, flags, payload=payload)
                r.sort_pairs(k, ko, kb, p, po, pb, ws, d, mode)
                assert np.array_equal(host(ko).view(bits.dtype), exp_k)
                assert np.array_equal(host(po), exp_p)
                r.sort_pairs(k, ko, kb, None, po, pb, ws, d, mode)   # payload = source index
                _, perm = oracle.radixsort(bits, flags, want_payload=True)
                assert np.array_equal(host(po), perm)


def test_indices_records(algo):
    """Records addressed by indices (the indirect config's shape): 32-byte records, key at +8."""
    rng = np.random.default_rng(21)
    n, stride = 200003, 32
    rec = rng.integers(0, 256, n * stride, dtype=np.uint8)
    keyview = rec.view(np.uint64)[1::4]
    keyview[:] = np.uint64(0xA5A5A5 << 40) | (keyview & np.uint64((1 << 12) - 1))  # top bytes constant
    idx = rng.permutation(n).astype(np.uint32)
    d_rec, d_idx = dev(rec), dev(idx)
    out = torch.empty_like(d_idx)
    for dt, d in ((np.uint64, 0), (np.int64, 3), (np.uint64, 2), (np.int64, 1)):
        kt = r.keytype_of(dt)
        ws = r.Workspace(kt, n, r.LAYOUT_INDIRECT)
        r.sort_indices(d_rec, stride, 8, d_idx, out, ws, dt, d, 64)
        exp = oracle.sort_indirect(rec, stride, 8, idx, dt, r.decode(kt, 64, d, True))
        assert np.array_equal(host(out), exp), (dt, d)
    # unaligned key offset (the packed 25-byte class of multiplatform/main.cpp:219-240, key at +9)
    n2, stride2 = 5001, 25
    rec2 = rng.integers(0, 256, n2 * stride2, dtype=np.uint8)
    idx2 = np.arange(n2, dtype=np.uint32)
    ws = r.Workspace(r.U64, n2, r.LAYOUT_INDIRECT)
    o2 = torch.empty(n2, dtype=torch.uint32, device="cuda")
    r.sort_indices(dev(rec2), stride2, 9, dev(idx2), o2, ws, np.uint64, 0, 64)
    exp = oracle.sort_indirect(rec2, stride2, 9, idx2, np.uint64, r.decode(r.U64, 64, 0, True))
    assert np.array_equal(host(o2), exp)


def test_pointer_indirection_levels(algo):
    """First-level (type+offset) and second-level (pointer member plus index) indirection, copy and
    in-place forms."""
    rng = np.random.default_rng(99)
    n = 30011
    rec = np.zeros((n, 4), np.uint64)                 # 32-byte records
    rec[:, 1] = rng.integers(0, 1 << 20, n, dtype=np.uint64)
    arr = rng.integers(-(1 << 31), 1 << 31, (n, 3), dtype=np.int64).astype(np.int32)  # second level
    d_rec, d_arr = dev(rec), dev(arr)
    rec[:, 2] = (d_arr.data_ptr() + np.arange(n, dtype=np.uint64) * 12).astype(np.uint64)
    d_rec = dev(rec)
    order = rng.permutation(n)
    ptrs = torch.from_numpy((d_rec.data_ptr() + order * 32).astype(np.int64)).cuda()
    ws = r.Workspace(r.U64, n, r.LAYOUT_INDIRECT)
    out = torch.empty_like(ptrs)
    for d in range(4):
        r.radixsortcopynoalloc_pointers(ptrs, out, ws, np.uint64, 8, d, 64)
        perm_rec = (host(out) - d_rec.data_ptr()) // 32
        exp = oracle.sort_indirect(rec.view(np.uint8).ravel(), 32, 8, order.astype(np.uint32), np.uint64,
                                   r.decode(r.U64, 64, d, True))
        assert np.array_equal(perm_rec, exp.astype(np.int64))
        # second level: key = ((int32 *)(rec->ptr_at_16))[1]
        r.radixsortcopynoalloc_pointers(ptrs, out, ws, np.int32, 16, d, 64, second_level=True, indirection2=0,
                                        index2=1)
        keys = arr[order, 1]
        _, perm = oracle.radixsort(keys, r.decode(r.I32, 64, d, True), want_payload=True)
        assert np.array_equal((host(out) - d_rec.data_ptr()) // 32, order[perm])
        # in place, both placements
        for move in (False, True):
            a, b = ptrs.clone(), torch.empty_like(ptrs)
            r.radixsortnoalloc_pointers(a, b, ws, np.uint64, 8, move, d, 64)
            assert np.array_equal((host(b if move else a) - d_rec.data_ptr()) // 32, exp.astype(np.int64))


def test_zero_and_one_element():
    ws = r.Workspace(r.U32, 1)
    a = torch.tensor([5], dtype=torch.int32, device="cuda")
    o = torch.zeros_like(a)
    r.radixsortcopynoalloc(a, o, torch.empty_like(a), ws)
    assert host(o).tolist() == [5]
    e = torch.empty(0, dtype=torch.int32, device="cuda")
    r.radixsortcopynoalloc(e, torch.empty_like(e), torch.empty_like(e), ws)
    assert r.radixsortcopy(a, o)


def test_workspace_reuse_many_calls(algo):
    """The look-back epoch advances per sort: many sorts of different sizes on one workspace."""
    rng = np.random.default_rng(5)
    ws = r.Workspace(r.U64, 1 << 18)
    for i in range(40):
        n = int(rng.integers(1, 1 << 18))
        bits = gen(8, n, "uniform" if i % 3 else "dups", rng)
        a = dev(bits)
        o, b = torch.empty_like(a), torch.empty_like(a)
        r.radixsortcopynoalloc(a, o, b, ws)
        assert np.array_equal(host(o), np.sort(bits))


@pytest.mark.parametrize("nbytes", [1, 2, 4, 8])
def test_small_pairs_one_lds_block(nbytes):
    """Sorts with a payload of up to 8192 keys run as one LDS block in the automatic plan: every
    shape boundary, every mode and direction, a given payload and the source index."""
    rng = np.random.default_rng(300 + nbytes)
    kt = KEYTYPES[nbytes][-1]
    for n in (2, 3, 100, 1023, 1024, 1025, 4096, 4097, 8191, 8192):
        bits = gen(nbytes, n, "dups", rng)
        payload = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
        ws = r.Workspace(kt, n, r.LAYOUT_PAIRS)
        k = dev(typed(bits, kt))
        ko, kb = torch.empty_like(k), torch.empty_like(k)
        p = dev(payload)
        po, pb = torch.empty_like(p), torch.empty_like(p)
        for mode in (64, 65, 66):
            for d in range(4):
                flags = r.decode(kt, mode, d, True)
                exp_k, exp_p = oracle.radixsort(bits, flags, payload=payload)
                r.sort_pairs(k, ko, kb, p, po, pb, ws, d, mode)
                assert np.array_equal(host(ko).view(bits.dtype), exp_k), (n, mode, d)
                assert np.array_equal(host(po), exp_p), (n, mode, d)
                r.sort_pairs(k, ko, kb, None, po, pb, ws, d, mode)
                _, perm = oracle.radixsort(bits, flags, want_payload=True)
                assert np.array_equal(host(po), perm), (n, mode, d)


def test_noalloc_rejects_host_arrays_on_device():
    """With a real device: a device workspace with a host input, or a device input with a host output,
    returns EINVAL before any launch; the Python forms refuse CPU tensors."""
    import ctypes
    L = r.lib()
    n = 4096
    ws = r.Workspace(r.U32, n)
    d = dev(np.arange(n, dtype=np.uint32))
    h = np.zeros(n, np.uint32)
    hp = h.ctypes.data_as(ctypes.c_void_p)
    dp = ctypes.c_void_p(d.data_ptr())
    o = torch.empty_like(d)
    op = ctypes.c_void_p(o.data_ptr())
    assert L.rsbd8_hip_radixsortcopynoalloc(r.U32, 0, n, ws.ptr(), ws.nbytes, hp, op, None, None) == -1
    assert L.rsbd8_hip_radixsortcopynoalloc(r.U32, 0, n, ws.ptr(), ws.nbytes, dp, hp, None, None) == -1
    assert L.rsbd8_hip_radixsortnoalloc(r.U32, 0, n, ws.ptr(), ws.nbytes, hp, op, 0, None) == -1
    assert not h.any()
    with pytest.raises(r.Rsbd8Error):
        r.radixsortcopynoalloc(d.cpu(), o, torch.empty_like(o), ws)
    # the same call with device arrays works
    r.radixsortcopynoalloc(d, o, torch.empty_like(o), ws)
    assert np.array_equal(host(o), np.arange(n, dtype=np.uint32))
