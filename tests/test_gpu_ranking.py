"""GPU parity of the stable in-tile ranking of the LSD scatter passes under every digit mix a wave
can hold. A wave finds the lanes that share its digit in one of three ways: the whole wave when
every lane holds lane 0's digit, eight ballots when at least 16 lanes do, and otherwise one LDS
word per digit that every lane ORs its bit into. The keys here are built in groups of 64 so that
the first pass sees each case (and both sides of the 16-lane threshold); later passes see the
mixes the permutation leaves. Sorted with the source index as payload, so the tie order is
checked bit for bit against the oracle.
"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
r = pytest.importorskip("radixsortbidi8_amd")


def regime_keys(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    groups = (n + 63) // 64
    out = np.empty((groups, 64, 4), np.uint8)
    kinds = rng.integers(0, 7, (groups, 4))
    for b in range(4):
        for k in range(7):
            sel = kinds[:, b] == k
            m = int(sel.sum())
            if m == 0:
                continue
            if k == 0:    # every lane the same digit
                v = np.repeat(rng.integers(0, 256, (m, 1)), 64, axis=1)
            elif k == 1:  # two digits
                v = rng.choice(np.array([17, 200], np.uint8), (m, 64))
            elif k == 2:  # four digits
                v = rng.integers(0, 4, (m, 64)) * 60
            elif k in (3, 4):  # lane 0's digit held by 16 or 15 lanes in all, the rest distinct
                share = 16 if k == 3 else 15
                v = np.empty((m, 64), np.int64)
                for i in range(m):
                    others = rng.permutation(np.arange(1, 256))[:64 - share]
                    row = np.concatenate([np.zeros(share, np.int64), others])
                    rest = rng.permutation(row[1:])
                    v[i] = np.concatenate([[0], rest]) + rng.integers(0, 256)
                v %= 256
            elif k == 5:  # uniform
                v = rng.integers(0, 256, (m, 64))
            else:         # lane 0 alone with its digit, everyone else shares another one
                v = np.full((m, 64), 9)
                v[:, 0] = 10
            out[sel, :, b] = v.astype(np.uint8)
    return out.reshape(-1, 4)[:n].copy().view("<u4").reshape(-1)


@pytest.mark.parametrize("algo", [1, 2], ids=["onesweep", "countscan"])
@pytest.mark.parametrize("cfg", [0, 2, 3])
def test_rank_regimes_with_index_payload(algo, cfg):
    n = (1 << 20) + 777
    keys = regime_keys(n, 40 + cfg)
    d_in = torch.from_numpy(keys.view(np.int32)).cuda().view(torch.uint32)
    ko, kb = torch.empty_like(d_in), torch.empty_like(d_in)
    po = torch.empty(n, dtype=torch.int32, device="cuda")
    pb = torch.empty_like(po)
    ws = r.Workspace(r.U32, n, r.LAYOUT_PAIRS)
    prev_a, prev_c = r.set_algorithm(algo), r.set_tile_config(cfg)
    try:
        for d in (0, 3):
            flags = r.decode(r.U32, 64, d, True)
            exp_k, exp_p = oracle.radixsort(keys, flags, want_payload=True)
            r.sort_pairs(d_in, ko, kb, None, po, pb, ws, d, 64)
            torch.cuda.synchronize()
            got_k = ko.view(torch.int32).cpu().numpy().view(np.uint32)
            got_p = po.cpu().numpy().view(np.uint32)
            assert np.array_equal(got_k, exp_k), f"keys differ (direction {d})"
            assert np.array_equal(got_p, exp_p), f"tie order differs (direction {d})"
    finally:
        r.set_algorithm(prev_a)
        r.set_tile_config(prev_c)


def test_regime_keys_cover_the_thresholds():
    """The generator really yields groups whose lane 0 digit is shared by exactly 16 and 15 lanes,
    by all 64, and by one."""
    k = regime_keys(64 * 4000, 3).view(np.uint8).reshape(-1, 64, 4)
    counts = set()
    for g in range(k.shape[0]):
        for b in range(4):
            counts.add(int((k[g, :, b] == k[g, 0, b]).sum()))
    assert {1, 15, 16, 64} <= counts
