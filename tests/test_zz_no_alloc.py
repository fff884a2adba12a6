"""rsbd8_hip_radixsortcopynoalloc_static (the entry point that allocates nothing: its workspace is the
library's file-scope device array) against rsbd8_hip_radixsortcopynoalloc with a workspace
argument, through the C ABI: bit-exact on consecutive calls of falling sizes (every plan family:
the MSD plan, the LSD passes, the 16-bit count fill, the single LDS block), at the capacity limit,
and the rejection of a count above it before any launch."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
r = pytest.importorskip("radixsortbidi8_amd")

MAX = 1 << 24
TDT = {r.U8: torch.uint8, r.U16: torch.uint16, r.U32: torch.uint32, r.I32: torch.int32, r.F32: torch.float32,
       r.U64: torch.uint64, r.I64: torch.int64, r.F64: torch.float64}


def _lib():
    L = r.lib()
    f = L.rsbd8_hip_radixsortcopynoalloc_static
    f.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p]
    f.restype = ctypes.c_int
    return L


def _keys(kt, n, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    nbytes = torch.empty(0, dtype=TDT[kt]).element_size()
    raw = torch.randint(-(1 << 62), 1 << 62, (n,), dtype=torch.int64, device="cuda", generator=g)
    raw = raw ^ (raw << 1)  # every bit varies
    return raw.view(torch.uint8)[: n * nbytes].clone().view(TDT[kt])


def _static(L, kt, flags, a, o, b):
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return L.rsbd8_hip_radixsortcopynoalloc_static(kt, flags, a.numel(), ctypes.c_void_p(a.data_ptr()),
                                                   ctypes.c_void_p(o.data_ptr()),
                                                   None if b is None else ctypes.c_void_p(b.data_ptr()), s)


def test_static_equals_workspace_form_on_falling_sizes():
    L = _lib()
    # (key type, mode, direction, count), largest first: each call reuses the array the previous
    # one left behind
    seq = [(r.U64, 64, 0, MAX), (r.I32, 64, 3, 1 << 22), (r.F32, 65, 1, 3_000_001), (r.U16, 64, 0, 1 << 22),
           (r.F64, 64, 2, 100_003), (r.U32, 64, 0, 12_000), (r.U8, 64, 1, 5000), (r.I64, 66, 3, 777), (r.U64, 64, 0, 2)]
    for i, (kt, mode, d, n) in enumerate(seq):
        print(f"static call {i}: keytype {kt} mode {mode} direction {d} count {n}", flush=True)
        flags = r.decode(kt, mode, d)
        a = _keys(kt, n, 0x5EED0100 + i)
        o_static, o_ref = torch.empty_like(a), torch.empty_like(a)
        b = torch.empty_like(a)
        o_static.view(torch.uint8).zero_()
        rc = _static(L, kt, flags, a, o_static, b)
        assert rc == 0, L.rsbd8_hip_last_error()
        ws = r.Workspace(kt, n)
        r.radixsortcopynoalloc(a, o_ref, torch.empty_like(a), ws, d, mode)
        torch.cuda.synchronize()
        assert torch.equal(o_static.view(torch.uint8), o_ref.view(torch.uint8)), (kt, mode, d, n)
        del ws


def test_static_twice_then_smaller():
    """The same call twice, then a smaller one of the same key type: stale contents of the shared
    array never leak into a later result."""
    L = _lib()
    for n in (1 << 21, 1 << 21, 65537, 1 << 20):
        a = _keys(r.U64, n, n)
        o, b = torch.empty_like(a), torch.empty_like(a)
        assert _static(L, r.U64, 0, a, o, b) == 0, L.rsbd8_hip_last_error()
        ref = torch.sort(a.view(torch.int64) ^ torch.tensor(-(1 << 63), dtype=torch.int64, device="cuda")).values
        torch.cuda.synchronize()
        assert torch.equal(o.view(torch.int64) ^ torch.tensor(-(1 << 63), dtype=torch.int64, device="cuda"), ref), n


def test_static_rejects_above_capacity_and_host_arrays():
    L = _lib()
    n = MAX + 1
    a = torch.zeros(n, dtype=torch.int32, device="cuda")
    o = torch.full_like(a, 7)
    assert _static(L, r.I32, 2, a, o, torch.empty_like(a)) == -1
    assert b"RSBD8_HIP_STATIC_MAX_COUNT" in L.rsbd8_hip_last_error()
    torch.cuda.synchronize()
    assert int(o.min().item()) == 7  # nothing ran
    h = np.zeros(1000, np.uint32)
    d = torch.zeros(1000, dtype=torch.int32, device="cuda")
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    buf = torch.empty_like(d)
    rc = L.rsbd8_hip_radixsortcopynoalloc_static(r.U32, 0, 1000, h.ctypes.data_as(ctypes.c_void_p),
                                                 ctypes.c_void_p(d.data_ptr()), ctypes.c_void_p(buf.data_ptr()), s)
    assert rc == -1 and b"not device memory" in L.rsbd8_hip_last_error()
