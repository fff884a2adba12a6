"""Regression: the single-block LDS sort skips a digit only when every key of the block agrees on
it, not when every key of each wave's chunk does.

The block's keys sit wave-major (wave w holds the 64*KPT consecutive logical keys from w*64*KPT),
so inputs whose digits change only at chunk boundaries (arange, reversed arange, l // chunk, long
runs of equal bytes) hold digits that are constant inside each wave and differ between waves. The
automatic plan sends every keys-only sort of up to 16384 keys and every sort with a payload of up
to 8192 keys to this kernel. Bit-exact against the oracle, every direction.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
r = pytest.importorskip("radixsortbidi8_amd")

UINT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
KT = {1: r.U8, 2: r.U16, 4: r.U32, 8: r.U64}
TDT = {1: torch.uint8, 2: torch.uint16, 4: torch.uint32, 8: torch.uint64}
SIZES = (1000, 1024, 3000, 4096, 8192, 12000, 16384)


def inputs(nbytes, n):
    """Keys whose digits are constant within each 256- or 1024-key chunk (the wave chunks of the
    block shapes: 256x4 up to 1024 keys, 16 keys per thread above)."""
    l = np.arange(n, dtype=np.uint64)
    hi = np.uint64(np.iinfo(UINT[nbytes]).max)
    out = {
        "arange": l & hi,
        "reversed": (np.uint64(n - 1) - l) & hi,
        "chunk256": (l // np.uint64(256)) & hi,
        "chunk1024": (l // np.uint64(1024)) & hi,
        "chunk1024_desc": (np.uint64(n // 1024 + 1) - l // np.uint64(1024)) & hi,
        "runs_hi": ((l // np.uint64(300)) << np.uint64(8 * nbytes - 8)) & hi,
    }
    return {k: v.astype(UINT[nbytes]) for k, v in out.items()}


@pytest.mark.parametrize("nbytes", [1, 2, 4, 8])
def test_lds_block_chunk_uniform_digits_keys_only(nbytes):
    kt = KT[nbytes]
    for n in SIZES:
        ws = r.Workspace(kt, n)
        for name, bits in inputs(nbytes, n).items():
            a = torch.from_numpy(bits).cuda()
            o, b = torch.empty_like(a), torch.empty_like(a)
            for mode in (64, 2, 3):
                for d in range(4):
                    flags = r.decode(kt, mode, d)
                    r.radixsortcopynoalloc(a.view(TDT[nbytes]), o, b, ws, d, mode)
                    torch.cuda.synchronize()
                    exp, _ = oracle.radixsort(bits, flags)
                    assert np.array_equal(o.cpu().numpy(), exp), (n, name, mode, d)


@pytest.mark.parametrize("nbytes", [1, 2, 4, 8])
def test_lds_block_chunk_uniform_digits_payload(nbytes):
    kt = KT[nbytes]
    for n in (1000, 1024, 3000, 4096, 6000, 8192):
        ws = r.Workspace(kt, n, r.LAYOUT_PAIRS)
        for name, bits in inputs(nbytes, n).items():
            k = torch.from_numpy(bits).cuda()
            ko, kb = torch.empty_like(k), torch.empty_like(k)
            po = torch.empty(n, dtype=torch.uint32, device="cuda")
            pb = torch.empty_like(po)
            for mode in (64, 3):
                for d in range(4):
                    flags = r.decode(kt, mode, d, True)
                    r.sort_pairs(k, ko, kb, None, po, pb, ws, d, mode)
                    torch.cuda.synchronize()
                    exp_k, perm = oracle.radixsort(bits, flags, want_payload=True)
                    assert np.array_equal(ko.cpu().numpy(), exp_k), (n, name, mode, d)
                    assert np.array_equal(po.cpu().numpy(), perm), (n, name, mode, d)
