"""CPU checks of the C ABI boundary: the HIP library loads, exports every entry point that
include/rsbd8_hip.h declares, and its mode decoding equals the oracle's restatement of
H:59778-59904 (decoding makes no GPU call)."""
import os
import re

import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "rsbd8_hip.h")).read()
    return sorted(set(re.findall(r"\b(rsbd8_hip_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    import radixsortbidi8_amd as r
    lib = r.lib()
    declared = _declared()
    assert len(declared) >= 13
    for name in declared:
        assert hasattr(lib, name), name


def test_decode_matches_oracle():
    import radixsortbidi8_amd as r
    for kt in range(15):  # the ten narrow key types and the five wide ones
        for mode in list(range(8)) + [64, 65, 66]:
            for d in range(4):
                for ind in (False, True):
                    assert r.decode(kt, mode, d, ind) == oracle.decode(kt, mode, d, ind), (kt, mode, d, ind)
    # the 80-bit layouts decode as long double (signed floating point), the 128-bit integers as such
    for kt in (r.LD80, r.LD96, r.LD128):
        assert r.decode(kt, 64, 0) == r.decode(r.F64, 64, 0) and r.decode(kt, 65, 0) == r.decode(r.F64, 65, 0)
    assert r.decode(r.U128, 65, 0) == r.decode(r.U64, 65, 0) and r.decode(r.I128, 66, 0) == r.decode(r.I64, 66, 0)


def test_decode_rejects_bad_values():
    import radixsortbidi8_amd as r
    with pytest.raises(r.Rsbd8Error):
        r.decode(3, 67, 0)
    with pytest.raises(r.Rsbd8Error):
        r.decode(3, 64, 4)
    with pytest.raises(r.Rsbd8Error):
        r.decode(15, 64, 0)


def test_workspace_size_grows_with_count():
    import radixsortbidi8_amd as r
    L = r.lib()
    a = L.rsbd8_hip_workspace_size(r.U64, 1 << 10, r.LAYOUT_KEYS)
    b = L.rsbd8_hip_workspace_size(r.U64, 1 << 20, r.LAYOUT_KEYS)
    c = L.rsbd8_hip_workspace_size(r.U64, 1 << 20, r.LAYOUT_INDIRECT)
    assert 0 < a < b < c
    assert c - b >= (1 << 20) * (8 + 8 + 4 + 4)


def _host_ptr(a):
    import ctypes
    return a.ctypes.data_as(ctypes.c_void_p)


def test_noalloc_entry_points_reject_host_arrays():
    """Every noalloc entry point checks, on the host and before any launch, that its arrays are GPU
    memory: host (malloc'd) arrays return EINVAL instead of reaching a kernel."""
    import numpy as np

    import radixsortbidi8_amd as r
    L = r.lib()
    n = 1024
    a, b, c = (np.zeros(n, np.uint32) for _ in range(3))
    p, q = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    ws = np.zeros(L.rsbd8_hip_workspace_size(r.U32, n, r.LAYOUT_INDIRECT), np.uint8)
    W, H = _host_ptr(ws), ws.size
    calls = {
        "radixsortcopynoalloc": lambda: L.rsbd8_hip_radixsortcopynoalloc(r.U32, 0, n, W, H, _host_ptr(a), _host_ptr(b),
                                                                          _host_ptr(c), None),
        "radixsortnoalloc": lambda: L.rsbd8_hip_radixsortnoalloc(r.U32, 0, n, W, H, _host_ptr(a), _host_ptr(b), 0, None),
        "sort_pairs": lambda: L.rsbd8_hip_sort_pairs(r.U32, 0, n, W, H, _host_ptr(a), _host_ptr(b), _host_ptr(c), None,
                                                     _host_ptr(p), _host_ptr(q), None),
        "sort_indices": lambda: L.rsbd8_hip_sort_indices(r.U32, 0, n, W, H, _host_ptr(a), 4, 0, _host_ptr(p),
                                                         _host_ptr(q), None),
        "radixsortcopynoalloc_indirect": lambda: L.rsbd8_hip_radixsortcopynoalloc_indirect(
            r.U32, 16, n, W, H, _host_ptr(a), _host_ptr(b), 0, 0, 0, 0, None),
        "radixsortnoalloc_indirect": lambda: L.rsbd8_hip_radixsortnoalloc_indirect(
            r.U32, 16, n, W, H, _host_ptr(a), _host_ptr(b), 0, 0, 0, 0, 0, None),
    }
    for name, call in calls.items():
        assert call() == -1, name
        assert b"not device memory" in L.rsbd8_hip_last_error(), name
    assert not b.any() and not q.any()  # nothing was written


def test_python_wrappers_reject_host_and_strided_tensors():
    torch = pytest.importorskip("torch")
    import radixsortbidi8_amd as r
    x = torch.zeros(64, dtype=torch.int32)
    with pytest.raises(r.Rsbd8Error, match="GPU tensor"):
        r._ptr(x)
    assert r._ptr(x, device=False) is not None          # the allocating forms stage host tensors
    with pytest.raises(r.Rsbd8Error, match="contiguous"):
        r._ptr(torch.zeros(64, dtype=torch.int32)[::2], device=False)
