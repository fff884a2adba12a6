"""GPU parity of the 16-bit count fill: keys-only sorts of 2-byte keys in the
bijective modes from 2^22 keys (every mode but the absolute-value signed and float modes and float order,
whose ties keep source order) take a 65536-count histogram, a scan and a fill instead of two LSD
passes. Bit-exact against the oracle for every such mode and direction, over key sets with empty
stretches of the key space (the fill's bucket search), one value, and presorted input; in place;
and on a workspace that an LSD sort uses next (the fill borrows and clears the look-back rows).
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
r = pytest.importorskip("radixsortbidi8_amd")

NOT_BIJECTIVE = (3, 6, 7)  # explicit modes (flag bits ABS 1, SIGN 2, FLTP 4) whose ties keep source order


def keysets(n, rng):
    yield "uniform", rng.integers(0, 1 << 16, n, dtype=np.uint16)
    yield "two ends", rng.choice(np.array([0, 1, 65534, 65535], np.uint16), n)
    yield "sixteen", rng.integers(0, 16, n, dtype=np.uint16) * np.uint16(4099)
    yield "one value", np.full(n, 0x8001, np.uint16)
    yield "ascending", np.sort(rng.integers(0, 1 << 16, n, dtype=np.uint16))


@pytest.mark.parametrize("n", [(1 << 22) + 13, 1 << 23])
def test_fill16_matches_oracle(n):
    rng = np.random.default_rng(n & 0xFFFF)
    out = torch.empty(n, dtype=torch.int16, device="cuda")
    buf = torch.empty_like(out)
    for kt in (r.U16, r.I16):
        ws = r.Workspace(kt, n)
        modes = [64] + [m for m in range(8) if m not in NOT_BIJECTIVE]
        for name, bits in keysets(n, rng):
            d_in = torch.from_numpy(bits.view(np.int16)).cuda()
            dt = torch.uint16 if kt == r.U16 else torch.int16
            src = d_in.view(dt)
            for mode in modes:
                for d in range(4):
                    exp, _ = oracle.radixsort(bits, r.decode(kt, mode, d))
                    r.radixsortcopynoalloc(src, out.view(dt), buf.view(dt), ws, d, mode)
                    got = out.cpu().numpy().view(np.uint16)
                    assert np.array_equal(got, exp), (kt, name, mode, d)


def test_fill16_in_place_and_workspace_reuse():
    n = (1 << 22) + 4097
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 1 << 16, n, dtype=np.uint16)
    # in place through the allocating template
    t = torch.from_numpy(bits.view(np.int16)).cuda().view(torch.uint16)
    assert r.radixsort(t)
    assert np.array_equal(t.view(torch.int16).cpu().numpy().view(np.uint16), np.sort(bits))
    # the fill, then an LSD sort (absolute signed mode 3: not bijective) and a fill again on one workspace
    ws = r.Workspace(r.I16, n)
    src = torch.from_numpy(bits.view(np.int16)).cuda()
    out, buf = torch.empty_like(src), torch.empty_like(src)
    for mode in (64, 3, 64, 3):
        exp, _ = oracle.radixsort(bits, r.decode(r.I16, mode, 0))
        r.profile_begin()
        r.radixsortcopynoalloc(src, out, buf, ws, 0, mode)
        prof = r.profile_end()
        assert np.array_equal(out.cpu().numpy().view(np.uint16), exp), mode
        # the fill runs no scatter pass; the absolute mode takes the LSD passes
        assert (prof["scatter"][1] == 0) == (mode == 64), (mode, prof)


def test_fill16_noalloc_into_buffer_or_in_place():
    """radixsortnoalloc: with movetobuffer the fill writes the buffer and leaves the input alone;
    without it, the input itself."""
    n = (1 << 22) + 31
    rng = np.random.default_rng(17)
    bits = rng.integers(0, 1 << 16, n, dtype=np.uint16)
    ws = r.Workspace(r.U16, n)
    for d in (0, 3):
        exp, _ = oracle.radixsort(bits, r.decode(r.U16, 64, d))
        src = torch.from_numpy(bits.view(np.int16)).cuda().view(torch.uint16)
        buf = torch.empty_like(src)
        r.radixsortnoalloc(src, buf, ws, True, d)
        assert np.array_equal(buf.view(torch.int16).cpu().numpy().view(np.uint16), exp), d
        assert np.array_equal(src.view(torch.int16).cpu().numpy().view(np.uint16), bits), d
        r.radixsortnoalloc(src, buf, ws, False, d)
        assert np.array_equal(src.view(torch.int16).cpu().numpy().view(np.uint16), exp), d
