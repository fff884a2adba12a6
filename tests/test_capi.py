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
    for kt in range(10):
        for mode in list(range(8)) + [64, 65, 66]:
            for d in range(4):
                for ind in (False, True):
                    assert r.decode(kt, mode, d, ind) == oracle.decode(kt, mode, d, ind), (kt, mode, d, ind)


def test_decode_rejects_bad_values():
    import radixsortbidi8_amd as r
    with pytest.raises(r.Rsbd8Error):
        r.decode(3, 67, 0)
    with pytest.raises(r.Rsbd8Error):
        r.decode(3, 64, 4)
    with pytest.raises(r.Rsbd8Error):
        r.decode(11, 64, 0)


def test_workspace_size_grows_with_count():
    import radixsortbidi8_amd as r
    L = r.lib()
    a = L.rsbd8_hip_workspace_size(r.U64, 1 << 10, r.LAYOUT_KEYS)
    b = L.rsbd8_hip_workspace_size(r.U64, 1 << 20, r.LAYOUT_KEYS)
    c = L.rsbd8_hip_workspace_size(r.U64, 1 << 20, r.LAYOUT_INDIRECT)
    assert 0 < a < b < c
    assert c - b >= (1 << 20) * (8 + 8 + 4 + 4)
